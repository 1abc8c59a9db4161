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
#include "mm_fft_dev.h"
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

// Spectral contrast (SpectralDescriptors.swift:137-204): one thread per
// (frame, band) of S [F, T]. The mean of the q lowest and of the q highest
// magnitudes of the band equal those of the sorted band; they are kept in two
// register lists (q <= KMAX, inserted into only when a value beats the
// current q-th one) and summed in ascending order as the reference sums its
// sorted band. Larger q (n_fft above 4096) finds the q-th value by
// bisection over the float bit patterns (magnitudes are >= 0, so bit order is
// value order) and sums around it.
constexpr int kContrastKmax = 32;

// l[0..k) holds the k lowest values in descending order (HIGH false) or the
// k highest in ascending order (HIGH true), so l[0] is always the value a
// newcomer must beat; v beats it: drop l[0] and insert v in order. Every
// index is a compile-time constant (the lists stay in registers).
template <int KMAX, bool HIGH>
__device__ __forceinline__ void keep_extreme(float (&l)[KMAX], float v, int k) {
  const float edge = HIGH ? INFINITY : -INFINITY;   // stands for l[k]
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    if (i < k) {
      const float nx = (i + 1 < k) ? l[i + 1] : edge;
      const bool nx_first = HIGH ? (nx < v) : (nx > v);
      const bool v_here = (i == 0) || (HIGH ? (l[i] < v) : (l[i] > v));
      l[i] = nx_first ? nx : (v_here ? v : l[i]);
    }
  }
}

// Sum of the k lowest (HIGH false) or highest values of col[t * T], t < n.
template <bool HIGH>
__device__ float sum_extreme_bisect(const float* col, int T, int n, int k) {
  auto bits = [&](int t) { return __float_as_uint(col[(int64_t)t * T]); };
  unsigned lo = 0, hi = 0x7f800000u;
  while (lo < hi) {
    const unsigned mid = HIGH ? lo + (hi - lo + 1) / 2 : lo + (hi - lo) / 2;
    int c = 0;
    for (int t = 0; t < n; ++t) c += HIGH ? (bits(t) >= mid) : (bits(t) <= mid);
    if (HIGH) {
      if (c >= k) lo = mid;
      else hi = mid - 1;
    } else {
      if (c >= k) hi = mid;
      else lo = mid + 1;
    }
  }
  const float x = __uint_as_float(lo);   // the k-th value
  float s = 0.0f;
  int c = 0;
  for (int t = 0; t < n; ++t) {
    const float v = col[(int64_t)t * T];
    if (HIGH ? (v > x) : (v < x)) {
      s += v;
      ++c;
    }
  }
  return s + (float)(k - c) * x;
}

// Sums of the k lowest and k highest values of col[t * T], t < n, k <= KM,
// each in ascending order.
template <int KM>
__device__ __forceinline__ void band_extremes(const float* col, int T, int n, int k, float& valley, float& peak) {
  float lo[KM], hi[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    lo[i] = INFINITY;
    hi[i] = -INFINITY;
  }
  auto take = [&](float v) {
    if (v < lo[0]) keep_extreme<KM, false>(lo, v, k);
    if (v > hi[0]) keep_extreme<KM, true>(hi, v, k);
  };
  // four loads in flight before the (branchy) list updates
  int t = 0;
  for (; t + 4 <= n; t += 4) {
    const float v0 = col[(int64_t)t * T], v1 = col[(int64_t)(t + 1) * T];
    const float v2 = col[(int64_t)(t + 2) * T], v3 = col[(int64_t)(t + 3) * T];
    take(v0);
    take(v1);
    take(v2);
    take(v3);
  }
  for (; t < n; ++t) take(col[(int64_t)t * T]);
  valley = 0.0f;
  peak = 0.0f;
#pragma unroll
  for (int i = KM - 1; i >= 0; --i)
    if (i < k) valley += lo[i];
#pragma unroll
  for (int i = 0; i < KM; ++i)
    if (i < k) peak += hi[i];
}

__global__ void k_contrast(const float* __restrict__ S, int T, const ContrastBands* __restrict__ cb,
                           float* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (f >= T) return;
  const int e0 = cb->edge[b], n = cb->edge[b + 1] - e0;
  float y = 0.0f;
  if (n > 0) {
    const int k = cb->q[b];
    const float* col = S + (int64_t)e0 * T + f;
    float valley = 0.0f, peak = 0.0f;
    // k is the same for the whole block (one band): the smallest list that holds it
    if (k <= 2) {
      band_extremes<2>(col, T, n, k, valley, peak);
    } else if (k <= 4) {
      band_extremes<4>(col, T, n, k, valley, peak);
    } else if (k <= 8) {
      band_extremes<8>(col, T, n, k, valley, peak);
    } else if (k <= 16) {
      band_extremes<16>(col, T, n, k, valley, peak);
    } else if (k <= kContrastKmax) {
      band_extremes<kContrastKmax>(col, T, n, k, valley, peak);
    } else {
      valley = sum_extreme_bisect<false>(col, T, n, k);
      peak = sum_extreme_bisect<true>(col, T, n, k);
    }
    valley /= (float)k;
    peak /= (float)k;
    {
      // both dB values rounded before the difference, as the reference forms
      // them (an FMA would leave one product unrounded: equal bands not 0)
#pragma clang fp contract(off)
      const float pdb = 10.0f * log10f(fmaxf(peak, 1e-10f));
      const float vdb = 10.0f * log10f(fmaxf(valley, 1e-10f));
      y = pdb - vdb;
    }
  }
  out[(int64_t)b * T + f] = y;
}

hipError_t launch_contrast(const float* S, int T, int nb, const ContrastBands* d_cb, float* out, hipStream_t st) {
  if (T <= 0 || nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_contrast, dim3((unsigned)((T + 255) / 256), (unsigned)nb), dim3(256), 0, st, S, T, d_cb, out);
  return hipGetLastError();
}

// Fused form for one clip: each wave computes the FFT of one frame with the
// last pass in registers (the fused tile kernel's register-split FFT), forms
// |X[k]| in registers (lane l holds k = l + 64 q + BL r; lane 0 also the
// Nyquist bin) and reduces it with wave sums, so the magnitude spectrogram
// never reaches HBM. Rolloff finds the first bin whose cumulative sum reaches
// the threshold by a binary search over k (each probe one masked wave sum).
// Sums are per-lane then across lanes instead of in bin order; the
// descriptors agree with the two-pass kernel above to FP32 rounding.
template <int LOG2M, int NW>
__global__ __launch_bounds__(NW * 64) void k_desc_fused(TileArgs a, int kind, float srf, float nfftf, float param,
                                                       float* __restrict__ out) {
  using S = FftShape<LOG2M>;
  constexpr int M = S::M, QL = S::QL, RL = S::RL, BL = S::BL;
  extern __shared__ float lds_f[];
  const int wave = threadIdx.x >> 6, lane0 = threadIdx.x & 63;
  float* sf = lds_f + wave * (M + M / 16);
  const int t_begin = blockIdx.x * a.tiles_per_wg;
  const int t_end = min(a.n_tiles, t_begin + a.tiles_per_wg);
  const float2 pf[S::Q0][S::R0] = {};
  for (int tile = t_begin; tile < t_end; ++tile) {
    int lane = lane0;
    __asm__ __volatile__("" : "+v"(lane));
    const int clip = ldc(a.tile_clip, tile);
    const int fr = ldc(a.tile_f0, tile) + wave;
    if (fr >= ldc(a.clip_frames, clip)) continue;   // wave-uniform; the FFT uses only this wave's LDS
    float2 z[QL][RL];
    frame_fft_split<LOG2M, float>(a, clip, fr, sf, lane, a.window, a.twp, z, pf, false);
    float2 x[QL][RL];
    float nyq;
    real_split_regs<LOG2M>(z, a.twN, lane, x, nyq);
    float v[QL][RL];
#pragma unroll
    for (int q = 0; q < QL; ++q)
#pragma unroll
      for (int r = 0; r < RL; ++r) v[q][r] = 0.5f * sqrtf(fmaf(x[q][r].x, x[q][r].x, x[q][r].y * x[q][r].y));
    const float vn = (lane == 0) ? 0.5f * fabsf(nyq) : 0.0f;   // bin M
    auto freq = [&](int k) { return (float)k * srf / nfftf; };
    auto bin = [&](int q, int r) { return lane + 64 * q + BL * r; };
    float y = 0.0f;
    if (kind == DESC_CENTROID || kind == DESC_BANDWIDTH) {
      float ws = (lane == 0) ? freq(M) * vn : 0.0f, tot = vn;
#pragma unroll
      for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          ws = fmaf(freq(bin(q, r)), v[q][r], ws);
          tot += v[q][r];
        }
      ws = wave_sum(ws);
      tot = wave_sum(tot);
      const float c = tot > 0.0f ? ws / tot : 0.0f;
      if (kind == DESC_CENTROID) {
        y = c;
      } else {
        const bool p2 = (param == 2.0f);
        auto dev = [&](int k, float val) {
          const float d = fabsf(freq(k) - c);
          return val * (p2 ? d * d : powf(d, param));
        };
        float m = (lane == 0) ? dev(M, vn) : 0.0f;
#pragma unroll
        for (int q = 0; q < QL; ++q)
#pragma unroll
          for (int r = 0; r < RL; ++r) m += dev(bin(q, r), v[q][r]);
        m = wave_sum(m);
        y = tot > 0.0f ? (p2 ? sqrtf(m / tot) : powf(m / tot, 1.0f / param)) : 0.0f;
      }
    } else if (kind == DESC_ROLLOFF) {
      float tot = vn;
#pragma unroll
      for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) tot += v[q][r];
      tot = wave_sum(tot);
      if (tot > 0.0f) {
        const float thr = param * tot;
        auto cum = [&](int K) {   // sum of bins 0..K
          float c = (K >= M) ? vn : 0.0f;
#pragma unroll
          for (int q = 0; q < QL; ++q)
#pragma unroll
            for (int r = 0; r < RL; ++r) c += (bin(q, r) <= K) ? v[q][r] : 0.0f;
          return wave_sum(c);
        };
        int lo = 0, hi = M;
        while (lo < hi) {   // wave-uniform: every lane holds the same sums
          const int mid = (lo + hi) >> 1;
          if (cum(mid) >= thr) hi = mid;
          else lo = mid + 1;
        }
        y = (cum(lo) >= thr) ? freq(lo) : 0.0f;
      }
    } else {   // flatness: power 2, amin 1e-10
      const float amin = 1e-10f;
      float raw = vn, ls = 0.0f, as = 0.0f;
      if (lane == 0) {
        const float pn = fmaxf(vn * vn, amin);
        ls = logf(pn);
        as = pn;
      }
#pragma unroll
      for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const float pv = fmaxf(v[q][r] * v[q][r], amin);
          raw += v[q][r];
          ls += logf(pv);
          as += pv;
        }
      raw = wave_sum(raw);
      ls = wave_sum(ls);
      as = wave_sum(as);
      if (raw >= amin) {
        const float am = as / (float)(M + 1);
        const float gm = expf(ls / (float)(M + 1));
        y = am > amin ? gm / am : 0.0f;
      }
    }
    if (lane == 0) out[fr] = y;
  }
}

template <int LOG2M, int NW>
static hipError_t launch_desc_fused_t(const TileArgs& a, int kind, float srf, float nfftf, float param, float* out,
                                      int grid, hipStream_t st) {
  constexpr int M = 1 << LOG2M;
  const size_t lds = (size_t)NW * (M + M / 16) * sizeof(float);
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(k_desc_fused<LOG2M, NW>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds_limit_bytes());
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL((k_desc_fused<LOG2M, NW>), dim3(grid), dim3(NW * 64), lds, st, a, kind, srf, nfftf, param, out);
  return hipGetLastError();
}

int desc_fused_waves(int n_fft) { return (n_fft == 1024 || n_fft == 2048) ? 16 : n_fft == 4096 ? 8 : 0; }

hipError_t launch_desc_fused(const TileArgs& a, int kind, float srf, float nfftf, float param, float* out, int grid,
                             hipStream_t st) {
  switch (a.n_fft) {
    case 1024: return launch_desc_fused_t<9, 16>(a, kind, srf, nfftf, param, out, grid, st);
    case 2048: return launch_desc_fused_t<10, 16>(a, kind, srf, nfftf, param, out, grid, st);
    case 4096: return launch_desc_fused_t<11, 8>(a, kind, srf, nfftf, param, out, grid, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_descriptor(const float* S, int F, int T, int kind, float srf, float nfftf, float param,
                             float* out, hipStream_t st) {
  if (T <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_descriptor, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, st, S, F, T, kind, srf, nfftf,
                     param, out);
  return hipGetLastError();
}

}  // namespace mm
