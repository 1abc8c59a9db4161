// Spectral descriptors on gfx950: per-frame reductions of the magnitude
// spectrogram S [F, T] (row-major, as the fused STFT kernel writes it).
//
// Reference semantics followed (zakkeown/MetalMom):
//   centroid, bandwidth   Sources/MetalMomCore/Features/SpectralDescriptors.swift:25-111
//   rolloff               SpectralDescriptors.swift:228-270
//   flatness              SpectralDescriptors.swift:295-340
//   C ABI                 Sources/MetalMomBridge/Bridge.swift:621-813
//
// One thread per frame; bins are summed in increasing order in FP32, the
// order the reference accumulates them in. Consecutive threads read
// consecutive frames of a bin row, so every load is coalesced.
#include "mm_common.h"
#include "mm_kernels.h"

namespace mm {

__global__ void k_descriptor(const float* __restrict__ S, int F, int T, int kind, float srf, float nfftf, float param,
                             float* __restrict__ out) {
  // frequency of bin k: Float(k) * sr / n_fft, left to right (SpectralDescriptors.swift:31-35)
  auto freq = [&](int k) { return (float)k * srf / nfftf; };
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const float* col = S + t;
  float y = 0.0f;
  if (kind == DESC_CENTROID || kind == DESC_BANDWIDTH) {
    float ws = 0.0f, tot = 0.0f;
    for (int k = 0; k < F; ++k) {
      const float v = col[(int64_t)k * T];
      ws += freq(k) * v;
      tot += v;
    }
    const float c = tot > 0.0f ? ws / tot : 0.0f;
    if (kind == DESC_CENTROID) {
      y = c;
    } else {   // bandwidth, p = param
      float m = 0.0f;
      for (int k = 0; k < F; ++k) {
        const float v = col[(int64_t)k * T];
        m += v * powf(fabsf(freq(k) - c), param);
      }
      y = tot > 0.0f ? powf(m / tot, 1.0f / param) : 0.0f;
    }
  } else if (kind == DESC_ROLLOFF) {   // roll_percent = param
    float tot = 0.0f;
    for (int k = 0; k < F; ++k) tot += col[(int64_t)k * T];
    if (tot > 0.0f) {
      const float thr = param * tot;
      float cum = 0.0f;
      int kr = 0;
      for (int k = 0; k < F; ++k) {
        cum += col[(int64_t)k * T];
        if (cum >= thr) {
          kr = k;
          break;
        }
      }
      y = freq(kr);
    }
  } else {   // flatness: power 2, amin 1e-10 (the bridge's defaults)
    const float amin = 1e-10f;
    float raw = 0.0f;
    for (int k = 0; k < F; ++k) raw += col[(int64_t)k * T];
    if (raw >= amin) {
      float lsum = 0.0f, asum = 0.0f;
      for (int k = 0; k < F; ++k) {
        const float v = col[(int64_t)k * T];
        const float p = fmaxf(v * v, amin);
        lsum += logf(p);
        asum += p;
      }
      const float am = asum / (float)F;
      const float gm = expf(lsum / (float)F);
      y = am > amin ? gm / am : 0.0f;
    }
  }
  out[t] = y;
}

hipError_t launch_descriptor(const float* S, int F, int T, int kind, float srf, float nfftf, float param,
                             float* out, hipStream_t st) {
  if (T <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_descriptor, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, st, S, F, T, kind, srf, nfftf,
                     param, out);
  return hipGetLastError();
}

}  // namespace mm
