// Onset strength on gfx950: spectral flux of the mel dB spectrogram.
//
// Reference semantics followed (zakkeown/MetalMom):
//   OnsetDetection.onsetStrength    Sources/MetalMomCore/Rhythm/OnsetDetection.swift:23-128
//   C ABI                           Sources/MetalMomBridge/Bridge.swift:1319-1358
//
// Input: the dB mel spectrogram [n_bands, T] before the top_db floor and the
// whole-matrix maximum as an ordered key (k_db writes both). Each band
// value is floored at max - top_db on the fly (Scaling.swift:73-114), then
// out = max(0, S[b][i + lag] - S[b][i]) at output frame o = i + pad, summed
// over the bands in band order and divided by n_bands (aggregate) or kept
// per band; output frames without a difference are 0.
#include "mm_common.h"
#include "mm_kernels.h"

namespace mm {

__global__ void k_onset(const float* __restrict__ db, int n_bands, int T, const unsigned* __restrict__ max_key,
                        float top_db, int lag, int pad, int n_out, int aggregate, float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = aggregate ? (int64_t)n_out : (int64_t)n_out * n_bands;
  if (idx >= n) return;
  const int o = (int)(idx % n_out);
  const int i = o - pad;
  if (i < 0 || i >= T - lag) {
    out[idx] = 0.0f;
    return;
  }
  const float floor_db = (top_db > 0.0f) ? key_to_float(*max_key) - top_db : -INFINITY;
  if (aggregate) {
    float sum = 0.0f;
    for (int b = 0; b < n_bands; ++b) {
      const float* row = db + (int64_t)b * T;
      const float diff = fmaxf(row[i + lag], floor_db) - fmaxf(row[i], floor_db);
      sum += fmaxf(0.0f, diff);
    }
    out[idx] = sum / (float)n_bands;
  } else {
    const float* row = db + (idx / n_out) * T;
    out[idx] = fmaxf(0.0f, fmaxf(row[i + lag], floor_db) - fmaxf(row[i], floor_db));
  }
}

hipError_t launch_onset(const float* db, int n_bands, int T, const unsigned* max_key, float top_db, int lag, int pad,
                        int n_out, int aggregate, float* out, hipStream_t st) {
  const int64_t n = aggregate ? (int64_t)n_out : (int64_t)n_out * n_bands;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_onset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, db, n_bands, T, max_key, top_db,
                     lag, pad, n_out, aggregate, out);
  return hipGetLastError();
}

}  // namespace mm
