// PCM sample conversion on gfx950: interleaved little-endian PCM frames ->
// one float32 channel.
//
// Integer PCM is scaled by 2^-(bits-1) (unsigned 8-bit centred at 128), IEEE
// float is taken as is; a frame becomes the mean of its channels (mono) or
// its first channel, the channel-0 rule of AudioIO.swift:115-117, 126-128
// (zakkeown/MetalMom). Samples are widened to double, summed in channel order
// and divided there, then rounded once to float32, so the result equals the
// float64 mean rounded to float32 (integers of up to 32 bits and float32
// samples are exact in double). One thread per frame; HBM-bound (block_align
// bytes in, 4 bytes out per frame).
#include "mm_common.h"
#include "mm_kernels.h"

namespace mm {

__device__ __forceinline__ double pcm_sample(const unsigned char* p, int fmt) {
  switch (fmt) {
    case PCM_U8:
      return ((double)p[0] - 128.0) * (1.0 / 128.0);
    case PCM_S16:
      return (double)*reinterpret_cast<const short*>(p) * (1.0 / 32768.0);
    case PCM_S24: {
      int v = (int)p[0] | ((int)p[1] << 8) | ((int)(signed char)p[2] << 16);
      return (double)v * (1.0 / 8388608.0);
    }
    case PCM_S32:
      return (double)*reinterpret_cast<const int*>(p) * (1.0 / 2147483648.0);
    case PCM_F32:
      return (double)*reinterpret_cast<const float*>(p);
    default:  // PCM_F64
      return *reinterpret_cast<const double*>(p);
  }
}

__global__ void k_pcm_to_mono(const unsigned char* __restrict__ raw, int64_t frames, int channels, int fmt,
                              int mix, float* __restrict__ out) {
  const int bytes = pcm_bytes(fmt);
  const int64_t ba = (int64_t)bytes * channels;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < frames;
       f += (int64_t)gridDim.x * blockDim.x) {
    const unsigned char* p = raw + f * ba;
    double v = pcm_sample(p, fmt);
    if (mix && channels > 1) {
      for (int c = 1; c < channels; ++c) v += pcm_sample(p + c * bytes, fmt);
      v /= (double)channels;
    }
    out[f] = (float)v;
  }
}

hipError_t launch_pcm_to_mono(const void* raw, int64_t frames, int channels, int fmt, int mix, float* out,
                              int n_cu, hipStream_t st) {
  if (frames <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((frames + 255) / 256, (int64_t)std::max(n_cu, 1) * 16);
  hipLaunchKernelGGL(k_pcm_to_mono, dim3((unsigned)blocks), dim3(256), 0, st,
                     static_cast<const unsigned char*>(raw), frames, channels, fmt, mix, out);
  return hipGetLastError();
}

}  // namespace mm
