// Windowed-sinc sample-rate conversion on gfx950.
//
// Reference semantics followed (zakkeown/MetalMom):
//   Resample.resample     Sources/MetalMomCore/Audio/Resample.swift:12-96
//   Resample.besselI0     Sources/MetalMomCore/Audio/Resample.swift:100-112
//   C ABI mm_resample     Sources/MetalMomBridge/Bridge.swift:1087-1104
//
// Output sample i sits at input position p = i / ratio (ratio = target /
// source, both in double). It is scale * sum_j x[j] * w_j over j in
// [floor(p) - support + 1, floor(p) + support] clipped to the signal, where
// scale = min(1, ratio), support = ceil(64 / scale), and w_j interpolates the
// filter table linearly at |j - p| * scale * 512; taps whose table index
// reaches the last entry are skipped and an output whose weights sum to zero
// or less is 0. The table holds sinc(x) * Kaiser(beta 6, 64 zero crossings) at
// x = i / 512, computed in FP64 and rounded to FP32.
//
// Kernel: persistent, one workgroup of 1024 threads per CU. The workgroup
// copies the 128 KB table into LDS once, then takes chunks of 1024 consecutive
// outputs (one per thread); the chunk's input window (1023 / ratio + 2 support
// samples) is staged in LDS too when it fits in 24 KB, so every tap reads LDS.
// Tap positions are formed in FP64 exactly as the reference forms them
// (j - p is exact in double, times the filter scale), the weights and sums in
// FP32. The work per output is 2 * support taps of ~13 VALU operations and
// three LDS reads: compute-bound, not HBM-bound (8 bytes of HBM per output).
//
// Polyphase path: with g = gcd(source, target), output i sits (i * source
// mod target) / g steps of g / target past input floor(i * source / target),
// so the tap weights take only P = target / g distinct rows ("phases"). The
// filter is symmetric, so row P - ph is row ph reversed (bit for bit: the
// tap distances are the same integers), and rows 0 .. P/2 suffice. When they
// fit in LDS they are computed once per rate pair (k_poly_table, from the same
// FP32 table and interpolation, the tap distance taken exactly from integers)
// and each tap is then one LDS weight read, one LDS sample read and one FMA;
// the sums run over the same taps, split over two or four accumulators. With
// one phase (target divides source) the row is read with scalar loads and,
// for an even decimation factor, the samples as aligned float2 / float4;
// decimation by exactly 2 takes k_resample_dec2 (two outputs per lane).
// Pairs with 2 .. 4096 phases take the phase-major kernel (k_resample_phase:
// one phase per wave, all P rows in HBM read by scalar loads, the input
// window transposed in LDS). Otherwise the kernel above forms the weights per
// tap.
#include "mm_common.h"
#include "mm_kernels.h"

namespace mm {

namespace {

constexpr int kZeros = 64;
constexpr int kOver = 512;
constexpr int kTabLen = kZeros * kOver + 1;   // 32769
constexpr int kTabLds = 32772;                // padded to a 16-byte multiple
constexpr int kRsThreads = 1024;
constexpr int kWinCap = 6144;                 // staged input samples (24 KB)
constexpr int kWinCapPoly = 8192;             // staged input samples of the polyphase path (32 KB)
constexpr int64_t kPolyMaxFloats = 30000;     // polyphase weight rows in LDS (117 KB)
constexpr int kPhaseMaxFloats = 36864;        // transposed input window of the phase-major path (144 KB)
constexpr int kPhaseMaxRows = 4096;           // phases of the phase-major path (full weight table in HBM)

__device__ double bessel_i0(double x) {
  double sum = 1.0, term = 1.0;
  const double half = x / 2.0;
  for (int k = 1; k < 50; ++k) {
    term *= half / (double)k;
    const double t2 = term * term;
    sum += t2;
    if (t2 < sum * 1e-20) break;
  }
  return sum;
}

}  // namespace

__global__ void k_tab_resample(float* __restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTabLen) return;
  const double pi = 3.14159265358979323846;
  const double x = (double)i / (double)kOver;
  const double s = (i == 0) ? 1.0 : sin(pi * x) / (pi * x);
  const double t = x / (double)kZeros;
  const double k = (t >= 1.0) ? 0.0 : bessel_i0(6.0 * sqrt(1.0 - t * t)) / bessel_i0(6.0);
  T[i] = (float)(s * k);
}

__global__ __launch_bounds__(kRsThreads) void k_resample(const float* __restrict__ x, int64_t n,
                                                        float* __restrict__ y, int64_t n_out, double ratio,
                                                        double fscale, float scale, int support, int win_cap,
                                                        const float* __restrict__ table) {
  extern __shared__ float lds[];
  float* T = lds;
  float* W = lds + kTabLds;
  for (int i = threadIdx.x; i < kTabLen; i += kRsThreads) T[i] = table[i];
  const int64_t n_chunks = (n_out + kRsThreads - 1) / kRsThreads;
  for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const int64_t i0 = c * kRsThreads;
    const int64_t i1 = (i0 + kRsThreads < n_out ? i0 + kRsThreads : n_out) - 1;
    // input window of the chunk: floor(i / ratio) is non-decreasing in i
    int64_t lo = (int64_t)((double)i0 / ratio) - support + 1;
    lo = lo < 0 ? 0 : lo;
    int64_t hi = (int64_t)((double)i1 / ratio) + support;
    hi = hi > n - 1 ? n - 1 : hi;
    const int64_t span = hi - lo + 1;
    const bool staged = span <= win_cap;
    __syncthreads();   // the table is loaded; the previous chunk's reads of W are done
    if (staged)
      for (int k = threadIdx.x; k < span; k += kRsThreads) W[k] = x[lo + k];
    __syncthreads();
    const int64_t i = i0 + threadIdx.x;
    if (i <= i1) {
      const double p = (double)i / ratio;
      const int64_t k = (int64_t)p;
      const int64_t js = (k - support + 1 < 0) ? 0 : k - support + 1;
      const int64_t je = (k + support > n - 1) ? n - 1 : k + support;
      const float* src = staged ? W + (js - lo) : x + js;
      const int taps = (int)(je - js + 1);
      double jd = (double)js;   // j as a double (exact), so j - p rounds as the reference's does
      float sum = 0.0f, wsum = 0.0f;
      for (int m = 0; m < taps; ++m, jd += 1.0) {
        const double fp = fabs(jd - p) * fscale;
        const double fl = floor(fp);
        const int fi = (int)fl;
        if (fi < kTabLen - 1) {
          const float fr = (float)(fp - fl);
          const float w = T[fi] * (1.0f - fr) + T[fi + 1] * fr;
          sum += src[m] * w;
          wsum += w;
        }
      }
      y[i] = wsum > 0.0f ? sum * scale : 0.0f;
    }
  }
}

// Weight rows of the polyphase path: W[phase * stride + m] for the tap
// j = floor(p) - support + 1 + m of an output in that phase; the tap distance
// |j - p| * scale * 512 = |(m - support + 1) * target - phase * g| * 512 / D
// with D = max(source, target), split exactly into table index and fraction.
__global__ void k_poly_table(const float* __restrict__ table, int target, int g, int D, int support, int taps,
                             int stride, int rows, float* __restrict__ W) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * stride) return;
  const int ph = e / stride, m = e - ph * stride;
  float w = 0.0f;
  if (m < taps) {
    int64_t s = (int64_t)(m - support + 1) * target - (int64_t)ph * g;
    s = s < 0 ? -s : s;
    const int64_t N = s * kOver;
    const int64_t q = N / D, r = N - q * D;
    if (q < kTabLen - 1) {
      const float fr = (float)((double)r / (double)D);
      w = table[q] * (1.0f - fr) + table[q + 1] * fr;
    }
  }
  W[e] = w;
}

__global__ __launch_bounds__(kRsThreads) void k_resample_poly(const float* __restrict__ x, int64_t n,
                                                             float* __restrict__ y, int64_t n_out, int source,
                                                             int target, int g, int support, int taps, int stride,
                                                             int phases, int w_floats, int win_cap, float scale,
                                                             float wsum1,
                                                             const float* __restrict__ wtab) {
  extern __shared__ float lds[];
  float* Wt = lds;
  float* X = lds + ((w_floats + 3) & ~3);
  for (int i = threadIdx.x; i < w_floats; i += kRsThreads) Wt[i] = wtab[i];
  const int64_t n_chunks = (n_out + kRsThreads - 1) / kRsThreads;
  for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const int64_t i0 = c * kRsThreads;
    const int64_t i1 = (i0 + kRsThreads < n_out ? i0 + kRsThreads : n_out) - 1;
    int64_t lo = (i0 * source) / target - support + 1;
    lo = lo < 0 ? 0 : lo;
    int64_t hi = (i1 * source) / target + support;
    hi = hi > n - 1 ? n - 1 : hi;
    const int64_t span = hi - lo + 1;
    const bool staged = span <= win_cap;
    __syncthreads();   // the weights are loaded; the previous chunk's reads of X are done
    if (staged)
      for (int k = threadIdx.x; k < span; k += kRsThreads) X[k] = x[lo + k];
    __syncthreads();
    // one phase and every tap of every output of the chunk inside the signal:
    // the weight row is the same for all lanes and read with scalar loads
    // from global memory (wave-uniform addresses), all taps used
    const bool uniform_row = staged && target / g == 1 && (i0 * source) / target - support + 1 >= 0 &&
                             (i1 * source) / target + support <= n - 1;
    const int64_t i = i0 + threadIdx.x;
    if (uniform_row) {
      if (i <= i1) {
        // the output is a whole multiple D of the input rate and lo is not
        // clamped, so this lane's window starts at X[D * threadIdx.x]; with
        // D even (and taps a multiple of the width) the window is read as
        // aligned float2 or float4, so the lanes of a wave read one
        // contiguous run of LDS
        const int D = source / target;
        const int64_t k = (i * source) / target;
        const float* src = X + (k - support + 1 - lo);
        float sum;
        if ((D & 3) == 0 && (taps & 3) == 0) {
          const float4* s4 = reinterpret_cast<const float4*>(src);
          float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll 4
          for (int m = 0; m < taps; m += 4) {
            const float4 v = s4[m >> 2];
            a0 = fmaf(v.x, wtab[m], a0);
            a1 = fmaf(v.y, wtab[m + 1], a1);
            a2 = fmaf(v.z, wtab[m + 2], a2);
            a3 = fmaf(v.w, wtab[m + 3], a3);
          }
          sum = (a0 + a1) + (a2 + a3);
        } else if ((D & 1) == 0) {
          const float2* s2 = reinterpret_cast<const float2*>(src);
          float a0 = 0.0f, a1 = 0.0f;
#pragma unroll 8
          for (int m = 0; m < taps; m += 2) {
            const float2 v = s2[m >> 1];
            a0 = fmaf(v.x, wtab[m], a0);
            a1 = fmaf(v.y, wtab[m + 1], a1);
          }
          sum = a0 + a1;
        } else {
          sum = 0.0f;
#pragma unroll 16
          for (int m = 0; m < taps; ++m) sum = fmaf(src[m], wtab[m], sum);
        }
        y[i] = wsum1 > 0.0f ? sum * scale : 0.0f;
      }
      continue;
    }
    if (i <= i1) {
      const int64_t A = i * source;
      const int64_t k = A / target;
      const int ph = (int)((A - k * target) / g);
      const int64_t j0 = k - support + 1;   // the tap of m = 0
      const int m0 = j0 < 0 ? (int)-j0 : 0;
      const int m1 = (j0 + taps - 1 > n - 1) ? (int)(n - 1 - j0) : taps - 1;
      // phases above P/2 read the mirrored row backwards
      const bool mir = ph > phases / 2;
      const float* w = mir ? Wt + (phases - ph) * stride + (taps - 1) : Wt + ph * stride;
      const int dw = mir ? -1 : 1;
      float sum = 0.0f, wsum = 0.0f;
      if (staged && m0 == 0 && m1 == taps - 1) {
        // every tap inside the signal (all but the first and last few
        // outputs): two accumulator chains over tap pairs (taps is even),
        // unrolled so that many LDS reads are in flight
        const float* src = X + (j0 - lo);
        float s0 = 0.0f, s1 = 0.0f, w0 = 0.0f, w1 = 0.0f;
#pragma unroll 4
        for (int m = 0; m < taps; m += 2) {
          const float a = w[dw * m], b = w[dw * (m + 1)];
          s0 = fmaf(src[m], a, s0);
          s1 = fmaf(src[m + 1], b, s1);
          w0 += a;
          w1 += b;
        }
        sum = s0 + s1;
        wsum = w0 + w1;
      } else if (staged) {
        const float* src = X + (j0 - lo);
        for (int m = m0; m <= m1; ++m) {
          const float wm = w[dw * m];
          sum += src[m] * wm;
          wsum += wm;
        }
      } else {
        const float* src = x + j0;
        for (int m = m0; m <= m1; ++m) {
          const float wm = w[dw * m];
          sum += src[m] * wm;
          wsum += wm;
        }
      }
      y[i] = wsum > 0.0f ? sum * scale : 0.0f;
    }
  }
}

// One phase, decimation by exactly 2 (44.1 -> 22.05, 48 -> 24, 32 -> 16 kHz):
// output i = scale * sum_m w[m] x[2 i - support + 1 + m]. A workgroup takes
// chunks of 2048 outputs; lane t computes outputs i0 + 2t and i0 + 2t + 1,
// whose windows start at samples 4t and 4t + 2 of the chunk's staged window,
// so every aligned float4 read (one contiguous run of LDS per wave) serves
// four taps of each output. The weight row is wave-uniform (scalar loads).
// Chunks whose window leaves the signal, or that reach past the last output,
// sum the clipped taps exactly from HBM.
constexpr int kDec2Out = 2 * kRsThreads;
__global__ __launch_bounds__(kRsThreads) void k_resample_dec2(const float* __restrict__ x, int64_t n,
                                                             float* __restrict__ y, int64_t n_out, int support,
                                                             int taps, float scale, float wsum1,
                                                             const float* __restrict__ w) {
  extern __shared__ float lds[];
  float* X = lds;
  const int L = 2 * kDec2Out + taps;   // staged samples: the last lane reads up to 4 * 1023 + taps + 3
  const int64_t n_chunks = (n_out + kDec2Out - 1) / kDec2Out;
  const int t = threadIdx.x;
  for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const int64_t i0 = c * kDec2Out;
    const int64_t lo = 2 * i0 - support + 1;
    const bool interior = lo >= 0 && lo + L - 1 <= n - 1 && i0 + kDec2Out <= n_out;
    __syncthreads();   // the previous chunk's reads of X are done
    if (interior)
      for (int k = t; k < L; k += kRsThreads) X[k] = x[lo + k];
    __syncthreads();
    const int64_t i = i0 + 2 * t;
    if (interior) {
      const float4* s4 = reinterpret_cast<const float4*>(X + 4 * t);
      float4 v = s4[0];
      float p0 = v.x * w[0], p1 = v.y * w[1];
      p0 = fmaf(v.z, w[2], p0);
      p1 = fmaf(v.w, w[3], p1);
      float q0 = v.z * w[0], q1 = v.w * w[1];
      const int nq = taps >> 2;
#pragma unroll 4
      for (int k = 1; k < nq; ++k) {
        v = s4[k];
        const int m = 4 * k;
        p0 = fmaf(v.x, w[m], p0);
        p1 = fmaf(v.y, w[m + 1], p1);
        p0 = fmaf(v.z, w[m + 2], p0);
        p1 = fmaf(v.w, w[m + 3], p1);
        q0 = fmaf(v.x, w[m - 2], q0);
        q1 = fmaf(v.y, w[m - 1], q1);
        q0 = fmaf(v.z, w[m], q0);
        q1 = fmaf(v.w, w[m + 1], q1);
      }
      v = s4[nq];
      q0 = fmaf(v.x, w[taps - 2], q0);
      q1 = fmaf(v.y, w[taps - 1], q1);
      y[i] = wsum1 > 0.0f ? (p0 + p1) * scale : 0.0f;
      y[i + 1] = wsum1 > 0.0f ? (q0 + q1) * scale : 0.0f;
    } else {
      for (int o = 0; o < 2; ++o) {
        const int64_t io = i + o;
        if (io >= n_out) break;
        const int64_t j0 = 2 * io - support + 1;
        const int m0 = j0 < 0 ? (int)-j0 : 0;
        const int m1 = (j0 + taps - 1 > n - 1) ? (int)(n - 1 - j0) : taps - 1;
        float s = 0.0f, ws = 0.0f;
        for (int m = m0; m <= m1; ++m) {
          s += x[j0 + m] * w[m];
          ws += w[m];
        }
        y[io] = ws > 0.0f ? s * scale : 0.0f;
      }
    }
  }
}

// Sum of each full weight row, in tap order, after the P rows: the sign that
// the per-output weight sum has for every output whose taps all lie inside
// the signal.
__global__ void k_row_sums(float* __restrict__ W, int rows, int taps, int stride) {
  const int ph = blockIdx.x * blockDim.x + threadIdx.x;
  if (ph >= rows) return;
  const float* w = W + (int64_t)ph * stride;
  float s = 0.0f;
  for (int m = 0; m < taps; ++m) s += w[m];
  W[(int64_t)rows * stride + ph] = s;
}

// Phase-major form of the polyphase path. With P phases and input step
// Dp = source / g per P outputs, output i = P q + r has phase (r Dp) mod P
// and first tap j0 = q Dp + b_r - support + 1, b_r = floor(r Dp / P) < Dp.
// A workgroup takes 64 G consecutive q and every r; a wave takes one (r,
// r + 1, group of 64 q) task, so its two weight rows are wave-uniform
// (scalar loads from the full P-row table in HBM) and its 64 lanes read the
// staged window at a = 64 Dp gq + Dp lane + b_r + m, each sample feeding
// both phases (output r + 1's window starts b_{r+1} - b_r samples later). The window is stored transposed,
// a -> Xt[(a mod Dp) * LD + a div Dp], so those 64 reads are 64 consecutive
// words: no bank conflicts, for any Dp. Taps run in segments of constant
// a div Dp (column steps of LD words). Workgroups whose window leaves the
// signal sum the clipped taps exactly from HBM instead. kLD > 0 fixes LD at
// compile time (65: 64 lanes and taps <= Dp, e.g. 48 -> 22.05 kHz), so the
// column steps are immediate offsets rather than one address add per tap.
template <int kLD>
__global__ __launch_bounds__(kRsThreads) void k_resample_phase(const float* __restrict__ x, int64_t n,
                                                              float* __restrict__ y, int64_t n_out, int P, int Dp,
                                                              int support, int taps, int stride, int G, int LD_,
                                                              float scale, const float* __restrict__ wfull) {
  const int LD = kLD > 0 ? kLD : LD_;
  extern __shared__ float lds[];
  float* Xt = lds;
  const float* rowsum = wfull + (int64_t)P * stride;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nq = 64 * G;
  const int L = nq * Dp + taps - 1;   // window samples of one workgroup
  const int64_t q_total = (n_out + P - 1) / P;
  const int64_t n_blocks = (q_total + nq - 1) / nq;
  const int Ph = (P + 1) / 2;   // phase pairs
  const int tasks = Ph * G;
  const int sd = kRsThreads / Dp, sm = kRsThreads - sd * Dp;
  for (int64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const int64_t q0 = blk * nq;
    const int64_t lo = q0 * Dp - support + 1;
    const bool interior = lo >= 0 && lo + L - 1 <= n - 1;
    __syncthreads();   // the previous block's reads of Xt are done
    if (interior) {
      int kd = threadIdx.x / Dp, km = threadIdx.x - kd * Dp;
      for (int k = threadIdx.x; k < L; k += kRsThreads) {
        Xt[km * LD + kd] = x[lo + k];
        kd += sd;
        km += sm;
        if (km >= Dp) {
          km -= Dp;
          ++kd;
        }
      }
    }
    __syncthreads();
    for (int t = wave; t < tasks; t += kRsThreads / 64) {
      // phases r and r + 1 (r + 1 = P: r again, not stored): their windows
      // start dl = b1 - b samples apart, so each sample read feeds both
      const int r = 2 * (t % Ph), gq = t / Ph;
      const bool two = r + 1 < P;
      const int64_t q = q0 + 64 * gq + lane;
      const int64_t i = q * P + r;
      const int64_t rd = (int64_t)r * Dp, rd1 = two ? rd + Dp : rd;
      const int ph = (int)(rd % P), b = (int)(rd / P);
      const int ph1 = (int)(rd1 % P), dl = (int)(rd1 / P) - b;
      const float* w = wfull + (int64_t)ph * stride;
      const float* w1 = wfull + (int64_t)ph1 * stride;
      float sum, sum1;
      bool pos, pos1;
      if (interior) {
        const float* xl = Xt + 64 * gq + lane;
        float a0 = 0.0f, a1 = 0.0f, c0 = 0.0f, c1 = 0.0f;
        // window samples s = 0 .. taps + dl - 1 from b: tap s of output r
        // (s < taps), tap s - dl of output r + 1 (s >= dl)
        const int s_end = taps + dl;
        int s = 0, col = b, drow = 0;
        if (Dp == 1) {   // upsampling by a whole factor: b = dl = 0, the window is contiguous
#pragma unroll 8
          for (; s < taps; s += 2) {
            const float v0 = xl[s], v1 = xl[s + 1];
            a0 = fmaf(v0, w[s], a0);
            a1 = fmaf(v1, w[s + 1], a1);
            c0 = fmaf(v0, w1[s], c0);
            c1 = fmaf(v1, w1[s + 1], c1);
          }
        }
        while (s < s_end) {
          const int len = min(s_end - s, Dp - col);   // samples of constant a div Dp
          const float* p = xl + col * LD + drow;
          const int e1 = min(len, max(0, dl - s));     // [0, e1): output r only
          const int e2 = min(len, max(0, taps - s));   // [e1, e2): both; [e2, len): output r + 1 only
          int k = 0;
          for (; k < e1; ++k) a0 = fmaf(p[k * LD], w[s + k], a0);
#pragma unroll 8
          for (; k + 1 < e2; k += 2) {
            const float v0 = p[k * LD], v1 = p[(k + 1) * LD];
            a0 = fmaf(v0, w[s + k], a0);
            a1 = fmaf(v1, w[s + k + 1], a1);
            c0 = fmaf(v0, w1[s + k - dl], c0);
            c1 = fmaf(v1, w1[s + k + 1 - dl], c1);
          }
          if (k < e2) {
            const float v0 = p[k * LD];
            a0 = fmaf(v0, w[s + k], a0);
            c0 = fmaf(v0, w1[s + k - dl], c0);
            ++k;
          }
          for (; k < len; ++k) c0 = fmaf(p[k * LD], w1[s + k - dl], c0);
          s += len;
          col = 0;
          ++drow;
        }
        sum = a0 + a1;
        sum1 = c0 + c1;
        pos = rowsum[ph] > 0.0f;
        pos1 = rowsum[ph1] > 0.0f;
      } else {
        float sv[2] = {0.0f, 0.0f}, wv[2] = {0.0f, 0.0f};
        for (int o = 0; o < 2; ++o) {
          if (i + o >= n_out || (o == 1 && !two)) break;
          const float* wo = o ? w1 : w;
          const int64_t j0 = q * Dp + b + (o ? dl : 0) - support + 1;
          const int m0 = j0 < 0 ? (int)-j0 : 0;
          const int m1 = (j0 + taps - 1 > n - 1) ? (int)(n - 1 - j0) : taps - 1;
          float s_ = 0.0f, ws = 0.0f;
          for (int m = m0; m <= m1; ++m) {
            s_ += x[j0 + m] * wo[m];
            ws += wo[m];
          }
          sv[o] = s_;
          wv[o] = ws;
        }
        sum = sv[0];
        sum1 = sv[1];
        pos = wv[0] > 0.0f;
        pos1 = wv[1] > 0.0f;
      }
      if (i < n_out) y[i] = pos ? sum * scale : 0.0f;
      if (two && i + 1 < n_out) y[i + 1] = pos1 ? sum1 * scale : 0.0f;
    }
  }
}

hipError_t launch_tab_resample(float* table, hipStream_t st) {
  hipLaunchKernelGGL(k_tab_resample, dim3((kTabLen + 255) / 256), dim3(256), 0, st, table);
  return hipGetLastError();
}

int resample_table_len() { return kTabLen; }

namespace {
int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    const int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}
}  // namespace

bool resample_poly_shape(int source_sr, int target_sr, int* phases, int* taps, int* stride) {
  const double ratio = (double)target_sr / (double)source_sr;
  const double scale = ratio < 1.0 ? ratio : 1.0;
  const int support = (int)std::ceil((double)kZeros / scale);
  const int64_t g = gcd64(source_sr, target_sr);
  const int64_t P = target_sr / g;
  const int t = 2 * support, st = t | 1;   // odd row stride: rows of different phases fall in different banks
  const int64_t rows = P / 2 + 1;          // rows 0 .. P/2; the others are their mirror images
  *phases = (int)std::min<int64_t>(P, 1 << 30);
  *taps = t;
  *stride = st;
  // the weights plus at least a 1024-output window must fit in LDS
  return rows * st <= kPolyMaxFloats && (int64_t)source_sr * 1023 / target_sr + t + 1 <= kWinCapPoly &&
         (size_t)(rows * st + kWinCapPoly + 4) * sizeof(float) <= lds_limit_bytes();
}

hipError_t launch_poly_table(int source_sr, int target_sr, const float* table, float* wtab, hipStream_t st) {
  int P, taps, stride;
  if (!resample_poly_shape(source_sr, target_sr, &P, &taps, &stride)) return hipErrorInvalidConfiguration;
  const int g = (int)gcd64(source_sr, target_sr);
  const int D = std::max(source_sr, target_sr);
  const int rows = P / 2 + 1;
  const int wf = rows * stride;
  hipLaunchKernelGGL(k_poly_table, dim3((wf + 255) / 256), dim3(256), 0, st, table, target_sr, g, D, taps / 2, taps,
                     stride, rows, wtab);
  return hipGetLastError();
}

hipError_t launch_resample_poly(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                const float* wtab, float wsum1, int grid, hipStream_t st) {
  int P, taps, stride;
  if (!resample_poly_shape(source_sr, target_sr, &P, &taps, &stride)) return hipErrorInvalidConfiguration;
  const int g = (int)gcd64(source_sr, target_sr);
  const int support = taps / 2;
  const int wf = (P / 2 + 1) * stride;
  const int64_t n_chunks = (n_out + kRsThreads - 1) / kRsThreads;
  if (n_chunks <= 0) return hipSuccess;
  // two workgroups per CU when both fit in LDS (one phase: 33 KB each), so
  // each SIMD holds 8 waves rather than 4 while the other waits on LDS
  const size_t lds = (size_t)(((wf + 3) & ~3) + kWinCapPoly) * sizeof(float);
  const int per_cu = 2 * lds <= lds_limit_bytes() ? 2 : 1;
  const int gr = (int)std::min<int64_t>((int64_t)std::max(grid, 1) * per_cu, n_chunks);
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(k_resample_poly),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds_limit_bytes());
  if (attr != hipSuccess) return attr;
  const double ratio = (double)target_sr / (double)source_sr;
  const float scale = (float)(ratio < 1.0 ? ratio : 1.0);
  hipLaunchKernelGGL(k_resample_poly, dim3(gr), dim3(kRsThreads), lds, st, x, n, y, n_out, source_sr, target_sr,
                     g, support, taps, stride, P, wf, kWinCapPoly, scale, wsum1, wtab);
  return hipGetLastError();
}

hipError_t launch_resample_dec2(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                const float* wtab, float wsum1, int grid, hipStream_t st) {
  int P, taps, stride;
  if (source_sr != 2 * target_sr || !resample_poly_shape(source_sr, target_sr, &P, &taps, &stride) || P != 1 ||
      (taps & 3) != 0)
    return hipErrorInvalidConfiguration;
  const int64_t n_chunks = (n_out + kDec2Out - 1) / kDec2Out;
  if (n_chunks <= 0) return hipSuccess;
  const size_t lds = (size_t)(2 * kDec2Out + taps) * sizeof(float);
  if (lds > lds_limit_bytes()) return hipErrorInvalidConfiguration;
  const int per_cu = 2 * lds <= lds_limit_bytes() ? 2 : 1;
  const int gr = (int)std::min<int64_t>((int64_t)std::max(grid, 1) * per_cu, n_chunks);
  hipLaunchKernelGGL(k_resample_dec2, dim3(gr), dim3(kRsThreads), lds, st, x, n, y, n_out, taps / 2, taps,
                     0.5f, wsum1, wtab);
  return hipGetLastError();
}

bool resample_phase_shape(int source_sr, int target_sr, PhaseShape* sh) {
  const double ratio = (double)target_sr / (double)source_sr;
  const double scale = ratio < 1.0 ? ratio : 1.0;
  const int support = (int)std::ceil((double)kZeros / scale);
  const int64_t g = gcd64(source_sr, target_sr);
  const int64_t P = target_sr / g, Dp = source_sr / g;
  if (P < 2 || P > kPhaseMaxRows || Dp > kRsThreads) return false;
  const int taps = 2 * support;
  auto rows_of = [&](int G) {   // LD: rows of the transposed window, odd
    const int64_t L = 64 * (int64_t)G * Dp + taps - 1;
    return ((L + Dp - 1) / Dp) | 1;
  };
  // enough (phase pair, group) tasks for 32 waves, as far as the window fits
  const int64_t Ph = (P + 1) / 2;
  int G = (int)std::max<int64_t>(1, std::min<int64_t>(16, (32 + Ph - 1) / Ph));
  while (G > 1 && Dp * rows_of(G) > kPhaseMaxFloats) --G;
  const int64_t LD = rows_of(G);
  if (Dp * LD > kPhaseMaxFloats || (size_t)(Dp * LD) * sizeof(float) > lds_limit_bytes()) return false;
  sh->phases = (int)P;
  sh->step = (int)Dp;
  sh->support = support;
  sh->taps = taps;
  sh->stride = taps | 1;
  sh->groups = G;
  sh->ld = (int)LD;
  return true;
}

hipError_t launch_phase_table(int source_sr, int target_sr, const float* table, float* wfull, hipStream_t st) {
  PhaseShape sh;
  if (!resample_phase_shape(source_sr, target_sr, &sh)) return hipErrorInvalidConfiguration;
  const int g = (int)gcd64(source_sr, target_sr);
  const int D = std::max(source_sr, target_sr);
  const int wf = sh.phases * sh.stride;
  hipLaunchKernelGGL(k_poly_table, dim3((wf + 255) / 256), dim3(256), 0, st, table, target_sr, g, D, sh.support,
                     sh.taps, sh.stride, sh.phases, wfull);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_row_sums, dim3((sh.phases + 63) / 64), dim3(64), 0, st, wfull, sh.phases, sh.taps, sh.stride);
  return hipGetLastError();
}

hipError_t launch_resample_phase(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                 const float* wfull, int grid, hipStream_t st) {
  PhaseShape sh;
  if (!resample_phase_shape(source_sr, target_sr, &sh)) return hipErrorInvalidConfiguration;
  const int64_t nq = 64 * (int64_t)sh.groups;
  const int64_t n_blocks = ((n_out + sh.phases - 1) / sh.phases + nq - 1) / nq;
  if (n_blocks <= 0) return hipSuccess;
  const size_t lds = (size_t)sh.step * sh.ld * sizeof(float);
  const int per_cu = 2 * lds <= lds_limit_bytes() ? 2 : 1;
  const int gr = (int)std::min<int64_t>((int64_t)std::max(grid, 1) * per_cu, n_blocks);
  static const hipError_t attr0 = hipFuncSetAttribute(reinterpret_cast<const void*>(k_resample_phase<0>),
                                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                                      (int)lds_limit_bytes());
  static const hipError_t attr65 = hipFuncSetAttribute(reinterpret_cast<const void*>(k_resample_phase<65>),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)lds_limit_bytes());
  if (attr0 != hipSuccess) return attr0;
  if (attr65 != hipSuccess) return attr65;
  const double ratio = (double)target_sr / (double)source_sr;
  const float scale = (float)(ratio < 1.0 ? ratio : 1.0);
  if (sh.ld == 65)
    hipLaunchKernelGGL(k_resample_phase<65>, dim3(gr), dim3(kRsThreads), lds, st, x, n, y, n_out, sh.phases,
                       sh.step, sh.support, sh.taps, sh.stride, sh.groups, sh.ld, scale, wfull);
  else
    hipLaunchKernelGGL(k_resample_phase<0>, dim3(gr), dim3(kRsThreads), lds, st, x, n, y, n_out, sh.phases,
                       sh.step, sh.support, sh.taps, sh.stride, sh.groups, sh.ld, scale, wfull);
  return hipGetLastError();
}

hipError_t launch_resample(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                           const float* table, int grid, hipStream_t st) {
  const double ratio = (double)target_sr / (double)source_sr;
  const double scale = ratio < 1.0 ? ratio : 1.0;
  const int support = (int)std::ceil((double)kZeros / scale);
  const int64_t n_chunks = (n_out + kRsThreads - 1) / kRsThreads;
  if (n_chunks <= 0) return hipSuccess;
  const int g = (int)std::min<int64_t>(std::max(grid, 1), n_chunks);
  // the table must fit; the staged window takes what is left, up to 24 KB
  const int64_t room = (int64_t)lds_limit_bytes() / 4 - kTabLds;
  if (room < 0) return hipErrorInvalidConfiguration;
  const int win_cap = (int)std::min<int64_t>(room, kWinCap);
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(k_resample), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit_bytes());
  if (attr != hipSuccess) return attr;
  const size_t lds = (size_t)(kTabLds + win_cap) * sizeof(float);
  hipLaunchKernelGGL(k_resample, dim3(g), dim3(kRsThreads), lds, st, x, n, y, n_out, ratio, scale * (double)kOver,
                     (float)scale, support, win_cap, table);
  return hipGetLastError();
}

}  // namespace mm
