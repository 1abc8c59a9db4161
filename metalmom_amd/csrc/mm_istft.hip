// Inverse STFT on gfx950: per-frame inverse real FFT, synthesis window and
// overlap-add normalised by the summed squared window.
//
// Reference semantics followed (zakkeown/MetalMom):
//   STFT.inverse                          Sources/MetalMomCore/Spectral/STFT.swift:686-861
//   DC / Nyquist imaginary parts ignored  STFT.swift:760-768 (vDSP packed format)
//   true inverse DFT scaling (1/n_fft)    STFT.swift:770-804
//   window-sum normalisation (> 1e-10)    STFT.swift:712-718, 820-826
//   centre trim / length                  STFT.swift:832-858
//   C ABI                                 Sources/MetalMomBridge/Bridge.swift:144-221
//
// Inverse real FFT of X[0..M] (N = 2M): with E = X[k] + conj(X[M-k]) and
// O = (X[k] - conj(X[M-k])) W_N^-k, the M-point sequence
// Z[k] = (E + i O) / 2 has the inverse DFT z[n] = x[2n] + i x[2n+1]. The
// kernels run the forward FFT on conj(E + i O) and read
// x[2n] = Re y[n] / N, x[2n+1] = -Im y[n] / N.
#include "mm_common.h"
#include "mm_fft_dev.h"
#include "mm_kernels.h"

#include <algorithm>

namespace mm {

// conj(E + i O) for bin n: xa = X[n], xb = X[M - n] (X[M] for n = 0),
// twn = W_N^n. DC and Nyquist imaginary parts are dropped.
__device__ __forceinline__ float2 inverse_split_in(float2 xa, float2 xb, bool dc, float2 twn) {
  if (dc) {
    xa.y = 0.0f;
    xb.y = 0.0f;
  }
  const float2 E = make_float2(xa.x + xb.x, xa.y - xb.y);
  const float2 D = make_float2(xa.x - xb.x, xa.y + xb.y);
  const float2 O = cmul(D, make_float2(twn.x, -twn.y));   // D W_N^-n
  return make_float2(E.x - O.y, -(E.y + O.x));
}

// W_32^r = exp(-2 pi i r / 32), r < 16: the pass-0 twiddle W_N^(j + r M/16) of
// the fused kernel is W_N^j W_32^r, one table load per butterfly.
__device__ __forceinline__ float2 w32(int r) {
  constexpr float C[16] = {1.0f, 0.980785280403230449f, 0.923879532511286756f, 0.831469612302545237f,
                           0.707106781186547524f, 0.555570233019602225f, 0.382683432365089772f,
                           0.195090322016128268f, 0.0f, -0.195090322016128268f, -0.382683432365089772f,
                           -0.555570233019602225f, -0.707106781186547524f, -0.831469612302545237f,
                           -0.923879532511286756f, -0.980785280403230449f};
  return make_float2(C[r], -C[(r >= 8) ? r - 8 : 8 - r]);   // (cos, -sin); sin(2 pi r/32) = C[|r - 8|]
}

// Frames whose window covers buffer sample s: f_lo .. f_hi (empty if
// f_lo > f_hi), clipped to the T frames that exist.
__device__ __forceinline__ void frame_range(int s, int hop, int n_fft, int T, int& f_lo, int& f_hi) {
  if (hop > 0) {
    const int fh = s / hop;
    const int i0 = s - fh * hop;                   // offset in frame fh
    const int cnt = (n_fft - 1 - i0) / hop + 1;    // frames with offset < n_fft
    f_lo = max(0, fh - cnt + 1);
    f_hi = min(T - 1, fh);
  } else {   // every frame starts at 0
    f_lo = 0;
    f_hi = (s < n_fft) ? T - 1 : -1;
  }
}

// Summed squared window of sample s over frames f_lo .. f_hi, in increasing
// frame order as the reference accumulates it (STFT.swift:712-718).
__device__ __forceinline__ float window_sum(int s, int hop, int f_lo, int f_hi, const float* __restrict__ window) {
  float ws = 0.0f;
  for (int f = f_lo; f <= f_hi; ++f) {
    const float w = window[s - f * hop];
    ws += w * w;
  }
  return ws;
}

// Output sample s of the overlap-add (s < total), all frames from one
// source: frame(f, i) returns windowed sample i of frame f.
template <class Frame>
__device__ __forceinline__ float ola_sample(int s, int hop, int n_fft, int T, const float* __restrict__ window,
                                            Frame frame) {
  int f_lo, f_hi;
  frame_range(s, hop, n_fft, T, f_lo, f_hi);
  float acc = 0.0f;
  for (int f = f_lo; f <= f_hi; ++f) acc += frame(f, s - f * hop);
  const float ws = window_sum(s, hop, f_lo, f_hi, window);
  return (ws > 1e-10f) ? acc / ws : acc;
}

// ---------------------------------------------------------------------------
// fused tile kernel (n_fft 1024 / 2048 / 4096)
// ---------------------------------------------------------------------------
//
// One workgroup = NW waves; tile t = frames g0 = t NW .. g0 + NW - 1 and
// output samples [g0 hop, (g0 + NW) hop). A workgroup runs a contiguous
// range of tiles and carries, from one tile to the next, the partial sums
// of the C = n_fft - hop samples after its output range that its frames
// touch (summed in increasing frame order from 0, so adding the next tile's
// frames reproduces the reference's order exactly). Its first tile is
// primed with the H = ceil(n_fft / hop) - 1 frames before it. Every frame's
// column is read once (plus H per workgroup); with a row stride that is a
// multiple of 16 frames, a tile reads whole 128-byte lines.

// Column transfer of a tile: thread t always handles frame g + t % NW and
// rows k = t / NW + 64 i (NW-frame runs of a row are contiguous: 128 bytes
// at NW = 16), KPT rows per thread.
template <int LOG2M, int NW> struct Cols {
  static constexpr int M = 1 << LOG2M, NT = NW * 64;
  static constexpr int KPT = ((M + 1) * NW + NT - 1) / NT;

  // rows of frames g_first + fl (fl < count, 0 <= frame < T) into v; others 0
  __device__ __forceinline__ static void load(const IstftArgs& a, int g_first, int count, float2 (&v)[KPT]) {
    const int fl = threadIdx.x % NW, k0 = threadIdx.x / NW;
    const int f = g_first + fl;
    const bool ok = fl < count && f >= 0 && f < a.T;
    const float2* src = a.X + (int64_t)k0 * a.ld + (ok ? f : 0);
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = k0 + 64 * i;
      v[i] = (ok && k <= M) ? src[(int64_t)64 * i * a.ld] : make_float2(0.0f, 0.0f);
    }
  }
  // into the regions: region fl holds X[0..M] of its frame
  __device__ __forceinline__ static void store(unsigned char* lds, int region_bytes, const float2 (&v)[KPT]) {
    const int fl = threadIdx.x % NW, k0 = threadIdx.x / NW;
    float2* dst = reinterpret_cast<float2*>(lds + fl * region_bytes);
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = k0 + 64 * i;
      if (k <= M) dst[k] = v[i];
    }
  }
};

// Inverse real FFT of frame g_first + wave (if wave < count and the frame
// exists), whose column is in the wave's region: conj(E + i O) in pass-0
// order through the register-split Stockham FFT, then the windowed frame
// stored in the region. Ends with the frames published by a barrier.
template <int LOG2M, int NW>
__device__ __forceinline__ void istft_ffts(const IstftArgs& a, unsigned char* lds, int g_first, int count,
                                           int wave) {
  using S = FftShape<LOG2M>;
  static_assert(S::REG_SPLIT && S::NP16 > 0, "register-split FFT shapes only");
  constexpr int M = S::M, N = 2 * M, R0 = S::R0, B0 = S::B0, Q0 = S::Q0;
  constexpr int QL = S::QL, RL = S::RL, BL = S::BL;
  using L0 = SplitPass<M, R0, 1>;
  const int f = g_first + wave;
  if (wave < count && f >= 0 && f < a.T) {
    const int lane = opaque(threadIdx.x & 63);   // keeps lane-derived addresses inside the tile loop
    unsigned char* region = lds + wave * a.region_bytes;
    const float2* Xc = reinterpret_cast<const float2*>(region);
    float2 v[Q0][R0];
    static_assert(R0 == 16 && B0 * 32 == N, "pass-0 twiddles are W_N^j W_32^r");
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) {
        const float2 tj = ldg<float2>(a.twN, (unsigned)j * 8u);
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int n = j + r * B0;
          v[q][r] = inverse_split_in(Xc[n], Xc[M - n], n == 0, (r == 0) ? tj : cmul(tj, w32(r)));
        }
      }
      __builtin_amdgcn_sched_barrier(0);   // one butterfly's reads in flight at a time
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) Dft<R0>::run(v[q]);
    }
    float2 z[QL][RL];
    SplitChain<LOG2M, 1, L0, R0>::run(reinterpret_cast<float*>(region), a.twp, lane, v, z);
    wave_lds_fence();
    const float inv_n = 1.0f / (float)N;
    float2* fo = reinterpret_cast<float2*>(region);   // sample pairs (x[2n], x[2n+1])
#pragma unroll
    for (int q = 0; q < QL; ++q) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int n = lane + 64 * q + BL * r;
        const float2 w = ldg<float2>(a.window, (unsigned)n * 8u);
        fo[n] = make_float2(z[q][r].x * inv_n * w.x, -z[q][r].y * inv_n * w.y);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  lds_barrier();
}

template <int LOG2M, int NW>
__global__ __launch_bounds__(NW * 64) void k_istft_tile(IstftArgs a) {
  constexpr int N = 2 << LOG2M;
  constexpr int NT = NW * 64;
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int hop = a.hop, T = a.T;
  const int H = a.halo, C = a.carry;
  const int R = H + 1;                    // frames covering a sample: R for offsets j < jsplit, else R - 1
  const int jsplit = N - (R - 1) * hop;   // in (0, hop]
  float* carry = reinterpret_cast<float*>(lds + NW * a.region_bytes);   // [2][carry_stride]
  float* wst = carry + 2 * a.carry_stride;                               // [hop]
  const int t_begin = blockIdx.x * a.tiles_per_wg;
  const int t_end = min(t_begin + a.tiles_per_wg, a.n_tiles);
  // summed squared window of a sample at offset j of its hop block with all
  // of its frames present, in increasing frame order (decreasing m)
  for (int j = threadIdx.x; j < hop; j += NT) {
    const int cnt = (j < jsplit) ? R : R - 1;
    float ws = 0.0f;
    for (int m = cnt - 1; m >= 0; --m) {
      const float w = a.window[m * hop + j];
      ws += w * w;
    }
    wst[j] = ws;
  }
  // sample index p -> (hop block b, offset j), stepped by NT without division
  const int b0 = (int)threadIdx.x / hop, j0 = (int)threadIdx.x - b0 * hop;
  const int db = NT / hop, dj = NT - db * hop;
  auto region_at = [&](int r, int i) { return reinterpret_cast<const float*>(lds + r * a.region_bytes)[i]; };
  float2 pf[Cols<LOG2M, NW>::KPT];
  bool have_pf = false;
  for (int ti = t_begin; ti < t_end; ++ti) {
    const int g0 = (a.tile0 + ti) * NW;
    float* cin = carry + ((ti - t_begin) & 1) * a.carry_stride;
    float* cout = carry + (((ti - t_begin) & 1) ^ 1) * a.carry_stride;
    if (ti == t_begin) {
      // prime: partial sums of samples [g0 hop, g0 hop + C) over frames g0-H .. g0-1
      const bool any = (H > 0) && (g0 > 0) && (g0 - H < T);
      if (any) {
        Cols<LOG2M, NW>::load(a, g0 - H, H, pf);
        Cols<LOG2M, NW>::store(lds, a.region_bytes, pf);
        lds_barrier();
        istft_ffts<LOG2M, NW>(a, lds, g0 - H, H, wave);
      }
      for (int i = threadIdx.x; i < C; i += NT) {
        float acc = 0.0f;
        if (any) {
          const int s = g0 * hop + i;
          int f_lo, f_hi;
          frame_range(s, hop, N, T, f_lo, f_hi);
          f_lo = max(f_lo, g0 - H);
          f_hi = min(f_hi, g0 - 1);
          for (int f = f_lo; f <= f_hi; ++f) acc += region_at(f - (g0 - H), s - f * hop);
        }
        cin[i] = acc;
      }
      lds_barrier();   // carry written, regions free
    }
    if (g0 < T) {   // workgroup-uniform
      if (!have_pf) Cols<LOG2M, NW>::load(a, g0, NW, pf);
      Cols<LOG2M, NW>::store(lds, a.region_bytes, pf);
      lds_barrier();
      istft_ffts<LOG2M, NW>(a, lds, g0, NW, wave);
    }
    // the next tile's columns, in flight through this tile's overlap-add
    // (barriers wait for LDS operations only)
    have_pf = (ti + 1 < t_end) && (g0 + NW < T);
    if (have_pf) Cols<LOG2M, NW>::load(a, g0 + NW, NW, pf);
    // output samples s = (g0 + b) hop + j, b < NW, within the requested range;
    // frame g0 + b - m covers s at offset m hop + j
    {
      const int64_t s_lo = a.start, s_hi = min(a.start + a.out_len, a.total);
      int b = b0, j = j0;
      for (int p = threadIdx.x; p < NW * hop; p += NT) {
        const int64_t sl = (int64_t)(g0 + b) * hop + j;
        if (sl >= s_lo && sl < a.start + a.out_len) {
          float y = 0.0f;
          if (sl < s_hi) {
            const int cnt = (j < jsplit) ? R : R - 1;
            const int m_hi = min(cnt - 1, g0 + b);            // frame >= 0
            const int m_lo = max(0, g0 + b - (T - 1));        // frame <= T - 1
            float acc = (p < C) ? cin[p] : 0.0f;              // frames before the tile
            for (int m = min(m_hi, b); m >= m_lo; --m) acc += region_at(b - m, m * hop + j);
            float ws;
            if (m_hi == cnt - 1 && m_lo == 0) {
              ws = wst[j];
            } else {   // near the ends of the signal: the frames that exist
              ws = 0.0f;
              for (int m = m_hi; m >= m_lo; --m) {
                const float w = a.window[m * hop + j];
                ws += w * w;
              }
            }
            y = (ws > 1e-10f) ? acc / ws : acc;
          }
          a.out[sl - a.start] = y;
        }
        b += db;
        j += dj;
        if (j >= hop) {
          j -= hop;
          ++b;
        }
      }
    }
    // carry for the next tile: samples (g0 + NW + b) hop + j, p = b hop + j < C,
    // over this tile's frames g0 + NW + b - m (m > b)
    {
      int b = b0, j = j0;
      for (int p = threadIdx.x; p < C; p += NT) {
        const int cnt = (j < jsplit) ? R : R - 1;
        const int m_lo = max(b + 1, g0 + NW + b - (T - 1));
        float acc = 0.0f;
        for (int m = cnt - 1; m >= m_lo; --m) acc += region_at(NW + b - m, m * hop + j);
        cout[p] = acc;
        b += db;
        j += dj;
        if (j >= hop) {
          j -= hop;
          ++b;
        }
      }
    }
    lds_barrier();   // the next tile rewrites the regions and reads the carry
  }
}

// ---------------------------------------------------------------------------
// generic path: any power-of-two n_fft from 2 to 16384, any hop >= 0
// ---------------------------------------------------------------------------

// One workgroup per frame: radix-2 Stockham in LDS (ping-pong) on
// conj(E + i O), then the windowed frame into frames[f][0..n_fft).
__global__ __launch_bounds__(256) void k_generic_ifft(IstftArgs a, float* __restrict__ frames) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int f = blockIdx.x;
  if (f >= a.T) return;
  const int N = a.n_fft, M = N / 2;
  float2* A = reinterpret_cast<float2*>(lds);
  float2* B = A + M;
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    const float2 xa = a.X[(int64_t)n * a.ld + f];
    const float2 xb = a.X[(int64_t)(M - n) * a.ld + f];
    A[n] = inverse_split_in(xa, xb, n == 0, a.twN[n]);
  }
  __syncthreads();
  for (int Ns = 1; Ns < M; Ns <<= 1) {
    for (int j = threadIdx.x; j < M / 2; j += blockDim.x) {
      const int k = j % Ns;
      float2 u = A[j], v = A[j + M / 2];
      v = cmul(v, a.tw[k * (M / (2 * Ns))]);   // W_{2Ns}^k
      const int base = (j / Ns) * Ns * 2 + k;
      B[base] = cadd(u, v);
      B[base + Ns] = csub(u, v);
    }
    __syncthreads();
    float2* t = A;
    A = B;
    B = t;
  }
  const float inv_n = 1.0f / (float)N;
  float* out = frames + (int64_t)f * N;
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    const float2 y = A[n];
    out[2 * n] = y.x * inv_n * a.window[2 * n];
    out[2 * n + 1] = -y.y * inv_n * a.window[2 * n + 1];
  }
}

__global__ void k_istft_ola(IstftArgs a, const float* __restrict__ frames) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= a.out_len) return;
  const int64_t s = o + a.start;
  float y = 0.0f;
  if (s < a.total) {
    const int N = a.n_fft;
    y = ola_sample((int)s, a.hop, N, a.T, a.window,
                   [&](int fr, int i) { return frames[(int64_t)fr * N + i]; });
  }
  a.out[o] = y;
}

// ---------------------------------------------------------------------------
// host-side planning and launch
// ---------------------------------------------------------------------------

static inline unsigned nblk_i(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

bool istft_fused_plan(int n_fft, int hop, IstftPlan* p) {
  if (n_fft != 1024 && n_fft != 2048 && n_fft != 4096) return false;
  if (hop <= 0) return false;
  const int M = n_fft / 2;
  const int w = (M > 1024) ? 8 : 16;
  const int R = (n_fft + hop - 1) / hop;   // frames touching one sample
  if (R - 1 > w / 2) return false;         // priming would be a large share of short ranges
  int rb = std::max((M + 1) * 8, std::max((M + M / 16) * 4, n_fft * 4));
  rb = (rb + 15) / 16 * 16;
  const int carry = std::max(n_fft - hop, 0);
  const int cs = (carry + 3) / 4 * 4;
  const size_t total = (size_t)w * rb + (size_t)2 * cs * sizeof(float) + (size_t)hop * sizeof(float);
  if (total > lds_limit_bytes()) return false;
  p->nw = w;
  p->halo = R - 1;
  p->carry = carry;
  p->carry_stride = cs;
  p->region_bytes = rb;
  p->lds_bytes = total;
  return true;
}

template <int LOG2M, int NW>
static hipError_t launch_istft_tile_t(const IstftArgs& a, int grid, size_t lds, hipStream_t st) {
  auto fn = k_istft_tile<LOG2M, NW>;
  static bool attr_done = false;
  if (!attr_done) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit_bytes());
    if (e != hipSuccess) return e;
    attr_done = true;
  }
  hipLaunchKernelGGL(fn, dim3(grid), dim3(NW * 64), lds, st, a);
  return hipGetLastError();
}

hipError_t launch_istft_tile(const IstftArgs& a, int nw, int grid, size_t lds, hipStream_t st) {
  if (a.n_fft == 1024 && nw == 16) return launch_istft_tile_t<9, 16>(a, grid, lds, st);
  if (a.n_fft == 2048 && nw == 16) return launch_istft_tile_t<10, 16>(a, grid, lds, st);
  if (a.n_fft == 4096 && nw == 8) return launch_istft_tile_t<11, 8>(a, grid, lds, st);
  return hipErrorInvalidValue;
}

hipError_t launch_istft_generic(const IstftArgs& a, float* frames, hipStream_t st) {
  const int M = a.n_fft / 2;
  const size_t lds = (size_t)2 * M * sizeof(float2);
  static bool attr_done = false;
  if (!attr_done) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_generic_ifft),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit_bytes());
    if (e != hipSuccess) return e;
    attr_done = true;
  }
  if (a.T > 0) hipLaunchKernelGGL(k_generic_ifft, dim3(a.T), dim3(256), lds, st, a, frames);
  if (a.out_len > 0) hipLaunchKernelGGL(k_istft_ola, dim3(nblk_i(a.out_len, 256)), dim3(256), 0, st, a, frames);
  return hipGetLastError();
}

}  // namespace mm
