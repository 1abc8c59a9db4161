// Device-side building blocks shared by the HIP kernels (gfx950): complex
// helpers, constant-address and global loads, barriers, the radix 2..16 DFT
// butterflies, Stockham FFT passes (LDS and register-resident with half-size
// exchanges), frame loaders for the fused kernel, the real-FFT split, dB and
// wave reductions. Included by every .hip file of the library; all functions
// are inline or templates.
#pragma once

#include "mm_common.h"

#include <climits>
#include <type_traits>

namespace mm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
  return make_float2(fmaf(a.x, w.x, -a.y * w.y), fmaf(a.x, w.y, a.y * w.x));
}
// a * (-i)
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// Round-to-nearest-even float -> bf16 bits (NaN stays NaN); one hardware
// conversion on gfx950.
__host__ __device__ inline uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __bf16 h = (__bf16)f;
  uint16_t b;
  __builtin_memcpy(&b, &h, 2);
  return b;
#else
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
#endif
}
__host__ __device__ inline float bf2f(uint16_t h) {
  unsigned u = ((unsigned)h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ inline uint16_t f2h(float f) {
  _Float16 h = (_Float16)f;
  uint16_t b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}
__device__ inline float h2f(uint16_t b) {
  _Float16 h;
  __builtin_memcpy(&h, &b, 2);
  return (float)h;
}

// Keeps the compiler from moving LDS accesses across this point. Within one
// wavefront LDS instructions execute in issue order, so no hardware wait is
// needed for wave-private scratch.
__device__ __forceinline__ void wave_lds_fence() { __asm__ __volatile__("" ::: "memory"); }

// A copy of v the compiler cannot see through: computations derived from it
// stay after this point instead of being scheduled (and kept live) earlier.
__device__ __forceinline__ int opaque(int v) {
  __asm__ __volatile__("" : "+v"(v));
  return v;
}
// Same for a wave-uniform value (kept in a scalar register).
__device__ __forceinline__ int opaque_s(int v) {
  __asm__ __volatile__("" : "+s"(v));
  return v;
}

// Workgroup barrier that waits for this wave's LDS (and scalar) operations
// only: vector memory loads issued before it stay in flight (the next tile's
// samples are prefetched across it), unlike __syncthreads, which also drains
// the vector memory counter.
__device__ __forceinline__ void lds_barrier() {
  __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Read-only launch data (tile and clip tables, block lists, prefetch
// records): read through the constant address space, which promises the
// compiler that the kernel does not write it, so wave-uniform reads become
// scalar loads (waited on the scalar counter) instead of vector loads plus a
// full vector-memory wait.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}
// Global load at base + byte_off through a 32-bit vector offset. With a
// wave-uniform base this is the scalar-base addressing form of global loads:
// one offset register per lane and no 64-bit address arithmetic per load.
// Constant parts of an address belong in the base (scalar adds only).
template <typename T>
__device__ __forceinline__ T ldg(const void* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(static_cast<const char*>(base) + byte_off);
}

// ldg for memory known to be global (HBM): the explicit global address space
// keeps global loads for pointers whose address space the compiler cannot
// infer (e.g. after wave_uniform), which would otherwise become flat loads.
// T: a builtin or ext_vector type.
template <typename T>
__device__ __forceinline__ T ldg_global(const void* base, unsigned byte_off) {
  using GC = const __attribute__((address_space(1))) char*;
  using GT = const __attribute__((address_space(1))) T*;
  return *(GT)((GC)base + byte_off);
}

// A pointer that is equal in all active lanes of the wave, moved into scalar
// registers (the compiler cannot see that a per-wave value is uniform).
template <typename T>
__device__ __forceinline__ const T* wave_uniform(const T* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return reinterpret_cast<const T*>(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ TilePf ldc_pf(const TilePf* p, int64_t i) {
  TilePf r;
  r.base = ldc(&p[i].base, 0);
  r.mask = ldc(&p[i].mask, 0);
  r.pad = 0;
  return r;
}

// Two floats -> their 16-bit (bf16 / f16, round to nearest even) hi parts
// and the 16-bit rounding of what each hi part leaves (lo), packed in pairs
// (element 0 in the low half): d ~ hi + lo to about twice the 16-bit
// precision.
template <int PREC>
__device__ __forceinline__ void split16(float d0, float d1, unsigned& hi, unsigned& lo) {
  typedef float f2v_ __attribute__((ext_vector_type(2)));
  if constexpr (PREC == PREC_BF16) {
    typedef __bf16 b2v __attribute__((ext_vector_type(2)));
    const b2v h = __builtin_convertvector((f2v_){d0, d1}, b2v);
    hi = __builtin_bit_cast(unsigned, h);
    const float h0 = __uint_as_float(hi << 16), h1 = __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f2v_){d0 - h0, d1 - h1}, b2v));
  } else {
    typedef _Float16 h2v_ __attribute__((ext_vector_type(2)));
    const h2v_ h = __builtin_convertvector((f2v_){d0, d1}, h2v_);
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f2v_){d0 - (float)h.x, d1 - (float)h.y}, h2v_));
  }
}

// Waits for every vector memory operation of this wave (s_waitcnt vmcnt(0);
// expcnt and lgkmcnt left at their maxima). Placed where outstanding loads
// are needed anyway, it also tells the compiler's wait tracking that nothing
// is pending, so registers that held in-flight loads on some path can be
// reused later without a conservative full wait.
__device__ __forceinline__ void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// 16 wait states after the last MFMA that wrote acc, before any other
// instruction can read it. The compiler puts only 6 between a
// v_mfma_f32_16x16x32_{bf16,f16} and a VALU copy of its result (the
// loop-carried accumulator of a multi-round contraction), and such a copy
// has been measured to read the accumulator as it was before that MFMA now
// and then (one chunk of a mel band lost, differing from run to run). The
// asm operand ties the wait to acc: it comes after acc's producers and
// before its readers.
__device__ __forceinline__ void mfma_settle(f32x4& acc) { asm volatile("s_nop 7\n\ts_nop 7" : "+v"(acc)); }

// Padded LDS index: one spare complex slot per 16 keeps the strided radix-16
// accesses of a Stockham pass on distinct banks.
__device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }
// For a stride C that is a multiple of 16, pidx(i + C) = pidx(i) + C + C/16:
// accesses at such strides use pidx(i) plus a constant (an LDS immediate offset).

// Index of sample `i` of a clip of `len` samples after centre padding.
// Returns -1 for a zero (constant) pad sample.
__device__ __forceinline__ int64_t pad_index(int64_t i, int64_t len, int pad_mode) {
  if (i >= 0 && i < len) return i;
  if (pad_mode == PAD_CONSTANT || len <= 0) return -1;
  if (len == 1) return 0;
  // numpy.pad(mode="reflect"): mirror without repeating the edge sample,
  // periodic with period 2*(len-1) for pads longer than the signal.
  const int64_t period = 2 * (len - 1);
  int64_t m = i % period;
  if (m < 0) m += period;
  return (m < len) ? m : period - m;
}

// ---------------------------------------------------------------------------
// forward DFTs of size 2, 4, 8, 16 in registers (W = exp(-2 pi i / R))
// ---------------------------------------------------------------------------

template <int R> struct Dft;

template <> struct Dft<1> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <> struct Dft<2> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct Dft<4> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = make_float2(t1.x + t3.y, t1.y - t3.x);  // t1 - i t3
    v[3] = make_float2(t1.x - t3.y, t1.y + t3.x);  // t1 + i t3
  }
};

template <> struct Dft<8> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float r = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e);
    Dft<4>::run(o);
    // o[k] *= W8^k
    float2 o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
    float2 o2 = mul_mi(o[2]);
    float2 o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

template <> struct Dft<16> {
  // 16 = 4 x 4: x[n1*4 + n2]; Y[n2][k1] = DFT4_n1; *= W16^(n2 k1); X[k1 + 4 k2] = DFT4_n2
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c = 0.92387953251128675613f;  // cos(pi/8)
    constexpr float s = 0.38268343236508977173f;  // sin(pi/8)
    constexpr float r = 0.70710678118654752440f;
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      float2 t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<4>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    // twiddles W16^(n2*k1)
    y[1][1] = cmul(y[1][1], make_float2(c, -s));
    y[1][2] = make_float2(r * (y[1][2].x + y[1][2].y), r * (y[1][2].y - y[1][2].x));
    y[1][3] = cmul(y[1][3], make_float2(s, -c));
    y[2][1] = make_float2(r * (y[2][1].x + y[2][1].y), r * (y[2][1].y - y[2][1].x));
    y[2][2] = mul_mi(y[2][2]);
    y[2][3] = make_float2(r * (y[2][3].y - y[2][3].x), -r * (y[2][3].x + y[2][3].y));
    y[3][1] = cmul(y[3][1], make_float2(s, -c));
    y[3][2] = make_float2(r * (y[3][2].y - y[3][2].x), -r * (y[3][2].x + y[3][2].y));
    y[3][3] = cmul(y[3][3], make_float2(-c, s));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      Dft<4>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// ---------------------------------------------------------------------------
// wave-level Stockham FFT of M complex points held in LDS scratch
// ---------------------------------------------------------------------------

template <int LOG2M> struct FftShape {
  static constexpr int M = 1 << LOG2M;
  static constexpr int NP16 = LOG2M / 4;          // radix-16 passes
  static constexpr int TAIL = 1 << (LOG2M % 4);   // final radix (1 = none)
  static constexpr int R0 = (NP16 > 0) ? 16 : TAIL;
  static constexpr int B0 = M / R0;               // butterflies of pass 0
  static constexpr int Q0 = (B0 + 63) / 64;
  static constexpr int NPASS = NP16 + (TAIL > 1 ? 1 : 0);
  static constexpr int RL = (TAIL > 1) ? TAIL : 16;   // radix of the last pass
  static constexpr int BL = M / RL;                   // its butterflies (= its output stride)
  static constexpr int QL = BL / 64;                  // butterflies per lane in the last pass
  // the last pass can stay in registers (and pair k with M-k by a lane
  // permutation) when every lane owns whole butterflies of it
  static constexpr bool REG_SPLIT = (NPASS >= 2) && (BL % 64 == 0);
};

// Per-pass twiddle tables: pass p (p >= 1, NS = 16^p, radix R_p) owns
// (R_p - 1) * NS entries laid out [r-1][k] = W_{NS R_p}^{k r}, so one lane
// reaches all of its pass twiddles from one base address plus constants.
template <int M, int P> struct PassTw {
  static constexpr int LOG2M = __builtin_ctz(M);
  static constexpr int NP16 = LOG2M / 4;
  static constexpr int TAIL = 1 << (LOG2M % 4);
  static constexpr int R = (P <= NP16 - 1) ? 16 : TAIL;   // pass P's radix (P >= 1)
  static constexpr int NS = 1 << (4 * P);
  static constexpr int OFF = PassTw<M, P - 1>::OFF + PassTw<M, P - 1>::SIZE;
  static constexpr int SIZE = (R - 1) * NS;
};
template <int M> struct PassTw<M, 0> {
  static constexpr int OFF = 0;
  static constexpr int SIZE = 0;
};

// One Stockham pass reading and writing LDS. NS = product of earlier radices.
template <int M, int R, int NS>
__device__ __forceinline__ void stockham_pass(float2* scr, const float2* __restrict__ tw, int lane_in) {
  const int lane = opaque(lane_in);
  constexpr int B = M / R;
  constexpr int QB = (B + 63) / 64;
  float2 v[QB][R];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int j = lane + 64 * q;
    if ((B % 64 == 0) || j < B) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (B % 16 == 0) v[q][r] = scr[pidx(j) + r * (B + B / 16)];
        else v[q][r] = scr[pidx(j + r * B)];
      }
    }
  }
  wave_lds_fence();
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int j = lane + 64 * q;
    if ((B % 64 == 0) || j < B) {
      const unsigned k = (unsigned)j % NS;
      constexpr int P = __builtin_ctz(NS) / 4;
#pragma unroll
      for (int r = 1; r < R; ++r)
        v[q][r] = cmul(v[q][r], ldg<float2>(tw + PassTw<M, P>::OFF + (r - 1) * NS, k * 8u));
      Dft<R>::run(v[q]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (NS % 16 == 0) scr[pidx(base) + r * (NS + NS / 16)] = v[q][r];
        else scr[pidx(base + r * NS)] = v[q][r];
      }
    }
  }
  wave_lds_fence();
}

template <int M, int P, int NPASS16, int TAIL, int PEND>
struct PassChain {
  // passes P .. PEND-1 (P >= 1), pass p with NS = 16^p
  __device__ __forceinline__ static void run(float2* scr, const float2* tw, int lane) {
    if constexpr (P < PEND) {
      constexpr int NS = 1 << (4 * P);
      if constexpr (P < NPASS16) {
        stockham_pass<M, 16, NS>(scr, tw, lane);
      } else {
        stockham_pass<M, TAIL, NS>(scr, tw, lane);
      }
      PassChain<M, P + 1, NPASS16, TAIL, PEND>::run(scr, tw, lane);
    }
  }
};

// Loads frame `fr` of a clip (pass-0 inputs, windowed) into registers, runs
// pass 0 from registers and the remaining passes in LDS. On return scr holds
// Z[k] = DFT_M(z)[k] in natural order (padded index), z[n] = x[2n] + i x[2n+1].
// Interior frames load straight from global memory; frames that touch the
// centre padding are first staged into the scratch by a plain loop.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

template <typename TIn> struct SamplePair;
// (x[2n], x[2n+1]) at base + byte offset of a signal in global memory (see ldg)
typedef float f2v __attribute__((ext_vector_type(2)));
template <> struct SamplePair<float> {
  __device__ __forceinline__ static float2 load(const float* base, unsigned byte_off) {
    const f2v v = ldg_global<f2v>(base, byte_off);
    return make_float2(v.x, v.y);
  }
  __device__ __forceinline__ static float one(const float* base, unsigned byte_off) {
    return ldg_global<float>(base, byte_off);
  }
};
template <> struct SamplePair<_Float16> {
  __device__ __forceinline__ static float2 load(const _Float16* base, unsigned byte_off) {
    const h2v h = ldg_global<h2v>(base, byte_off);
    return make_float2((float)h.x, (float)h.y);
  }
  __device__ __forceinline__ static float one(const _Float16* base, unsigned byte_off) {
    return (float)ldg_global<_Float16>(base, byte_off);
  }
};

// Pass-0 inputs of one frame, windowed, in registers: v[q][r] = z[j + r*B0]
// with z[n] = (x[2n] w[2n], x[2n+1] w[2n+1]), j = lane + 64 q. Interior frames
// load straight from global memory (one 8-byte / 4-byte pair load per complex
// value when the frame start is even); frames that touch the centre padding
// are staged through the scratch by a plain loop.
template <int LOG2M, typename TIn>
__device__ __forceinline__ void frame_load(const TileArgs& a, const TIn* __restrict__ x, int64_t len, int64_t s0,
                                           float2* scr, int lane_in,
                                           float2 (&v)[FftShape<LOG2M>::Q0][FftShape<LOG2M>::R0]) {
  const int lane = opaque(lane_in);
  using S = FftShape<LOG2M>;
  constexpr int M = S::M, R0 = S::R0, B0 = S::B0, Q0 = S::Q0;
  const bool interior = (s0 >= 0) && (s0 + 2 * M <= len);
  // x + s0 is the same in every lane (one frame per wave): scalar base, and
  // each load adds only the lane's byte offset
  if (interior && ((s0 & 1) == 0)) {
    const TIn* xf = wave_uniform(x + s0);
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if ((B0 % 64 == 0) || j < B0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const float2 w = ldg<float2>(a.window + 2 * r * B0, (unsigned)j * 8u);
          const float2 xs = SamplePair<TIn>::load(xf + 2 * r * B0, (unsigned)j * (2u * sizeof(TIn)));
          v[q][r] = make_float2(xs.x * w.x, xs.y * w.y);
        }
      }
    }
  } else if (interior) {
    const TIn* xf = wave_uniform(x + s0);
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if ((B0 % 64 == 0) || j < B0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const float2 w = ldg<float2>(a.window + 2 * r * B0, (unsigned)j * 8u);
          const unsigned o = (unsigned)j * (2u * sizeof(TIn));
          v[q][r] = make_float2(SamplePair<TIn>::one(xf + 2 * r * B0, o) * w.x,
                                SamplePair<TIn>::one(xf + 2 * r * B0 + 1, o) * w.y);
        }
      }
    }
  } else {
#pragma unroll 1
    for (int n = lane; n < M; n += 64) {
      const int64_t i0 = pad_index(s0 + 2 * n, len, a.pad_mode);
      const int64_t i1 = pad_index(s0 + 2 * n + 1, len, a.pad_mode);
      const float x0 = (i0 >= 0) ? (float)x[i0] : 0.0f;
      const float x1 = (i1 >= 0) ? (float)x[i1] : 0.0f;
      scr[pidx(n)] = make_float2(x0 * a.window[2 * n], x1 * a.window[2 * n + 1]);
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if ((B0 % 64 == 0) || j < B0) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          if constexpr (B0 % 16 == 0) v[q][r] = scr[pidx(j) + r * (B0 + B0 / 16)];
          else v[q][r] = scr[pidx(j + r * B0)];
        }
      }
    }
    wave_lds_fence();
  }
}

// FFT of frame `fr`: on return scr holds Z = DFT_M(z) in natural order
// (padded index), z[n] = x[2n] + i x[2n+1] windowed.
template <int LOG2M, typename TIn>
__device__ __forceinline__ void frame_fft(const TileArgs& a, int clip, int fr, float2* scr, int lane) {
  using S = FftShape<LOG2M>;
  constexpr int M = S::M, R0 = S::R0, B0 = S::B0, Q0 = S::Q0;
  const int64_t len = ldc(a.clip_len, clip);
  const int64_t s0 = (int64_t)fr * a.hop - (a.center ? M : 0);
  float2 v[Q0][R0];
  frame_load<LOG2M>(a, static_cast<const TIn*>(a.signal) + ldc(a.clip_off, clip), len, s0, scr, lane, v);
  // pass 0 (NS = 1: no twiddles), written in Stockham order
#pragma unroll
  for (int q = 0; q < Q0; ++q) {
    const int j = lane + 64 * q;
    if ((B0 % 64 == 0) || j < B0) {
      Dft<R0>::run(v[q]);
#pragma unroll
      for (int r = 0; r < R0; ++r) {   // R0 = 16: pidx(16 j + r) = 17 j + r
        if constexpr (R0 == 16) scr[17 * j + r] = v[q][r];
        else scr[pidx(j * R0 + r)] = v[q][r];
      }
    }
  }
  wave_lds_fence();
  if constexpr (S::NP16 > 0) PassChain<M, 1, S::NP16, S::TAIL, S::NPASS>::run(scr, a.twp, lane);
}

// ---- register-resident Stockham FFT with half-size LDS exchanges ----
//
// Every pass computes on registers; between passes a wave moves its values
// from the writer's layout to the reader's layout through a float buffer of
// M + M/16 entries, real parts first and imaginary parts second (half the LDS
// of a float2 exchange; a wave's LDS operations run in order, so the second
// half may overwrite the first as soon as it has been read).
//
// Layouts of pass p (radix R, NS = product of earlier radices, B = M / R
// butterflies, lane j = lane + 64 q owns butterfly j when j < B):
//   read:  element j + r B        (r < R)
//   write: element (j / NS) NS R + j % NS + r NS
template <int M, int R, int NS> struct SplitPass {
  static constexpr int B = M / R;
  static constexpr int QB = (B + 63) / 64;
  static constexpr bool ALL = (B % 64 == 0);   // every lane owns butterflies
  __device__ __forceinline__ static bool ok(int j) { return ALL || j < B; }
  // padded float index of read element (j, r)
  __device__ __forceinline__ static int rd(int j, int r) {
    if constexpr (B % 16 == 0) return pidx(j) + r * (B + B / 16);
    else return pidx(j + r * B);
  }
  // padded float index of written element (j, r)
  __device__ __forceinline__ static int wr(int j, int r) {
    if constexpr (NS == 1 && R == 16) return 17 * j + r;   // pidx(16 j + r)
    else if constexpr (NS % 16 == 0) return pidx((j / NS) * NS * R + j % NS) + r * (NS + NS / 16);
    else return pidx((j / NS) * NS * R + j % NS + r * NS);
  }
};

// Values written in pass W's layout, read back in pass Rd's layout.
template <class W, class Rd, int RW, int RR>
__device__ __forceinline__ void exchange_split(float* sf, int lane, const float2 (&in)[W::QB][RW],
                                               float2 (&out)[Rd::QB][RR]) {
#pragma unroll
  for (int half = 0; half < 2; ++half) {
#pragma unroll
    for (int q = 0; q < W::QB; ++q) {
      const int j = lane + 64 * q;
      if (W::ok(j)) {
#pragma unroll
        for (int r = 0; r < RW; ++r) sf[W::wr(j, r)] = half ? in[q][r].y : in[q][r].x;
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < Rd::QB; ++q) {
      const int j = lane + 64 * q;
      if (Rd::ok(j)) {
#pragma unroll
        for (int r = 0; r < RR; ++r) {
          if (half) out[q][r].y = sf[Rd::rd(j, r)];
          else out[q][r].x = sf[Rd::rd(j, r)];
        }
      }
    }
    wave_lds_fence();
  }
}

// Twiddles and DFT of pass P on registers (P >= 1).
template <int M, int P, int R, int NS>
__device__ __forceinline__ void split_pass_compute(const float2* __restrict__ tw, int lane,
                                                   float2 (&v)[SplitPass<M, R, NS>::QB][R]) {
  using L = SplitPass<M, R, NS>;
#pragma unroll
  for (int q = 0; q < L::QB; ++q) {
    const int j = lane + 64 * q;
    if (L::ok(j)) {
      const unsigned k = (unsigned)j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r)
        v[q][r] = cmul(v[q][r], ldg<float2>(tw + PassTw<M, P>::OFF + (r - 1) * NS, k * 8u));
      Dft<R>::run(v[q]);
    }
  }
}

// Passes P .. NPASS-1 given pass P-1's outputs `in` (in pass P-1's layout
// WL with radix RW); the last pass's outputs land in z.
template <int LOG2M, int P, class WL, int RW>
struct SplitChain {
  using S = FftShape<LOG2M>;
  static constexpr int M = S::M;
  static constexpr bool LAST = (P == S::NPASS - 1);
  static constexpr int R = (P < S::NP16) ? 16 : S::TAIL;
  static constexpr int NS = 1 << (4 * P);
  using L = SplitPass<M, R, NS>;
  __device__ __forceinline__ static void run(float* sf, const float2* tw, int lane, const float2 (&in)[WL::QB][RW],
                                             float2 (&z)[S::QL][S::RL]) {
    if constexpr (LAST) {
      static_assert(R == S::RL && L::QB == S::QL, "last pass shape");
      exchange_split<WL, L, RW, R>(sf, lane, in, z);
      split_pass_compute<M, P, R, NS>(tw, lane, z);
    } else {
      float2 v[L::QB][R];
      exchange_split<WL, L, RW, R>(sf, lane, in, v);
      split_pass_compute<M, P, R, NS>(tw, lane, v);
      SplitChain<LOG2M, P + 1, L, R>::run(sf, tw, lane, v, z);
    }
  }
};

// Pass-0 inputs of a frame that touches the centre padding: staged through
// the float buffer in two halves (x[2n] terms, then x[2n+1] terms).
template <int LOG2M, typename TIn>
__device__ __forceinline__ void frame_load_padded_split(const TileArgs& a, const TIn* __restrict__ x, int64_t len,
                                                        int64_t s0, float* sf, int lane, const float* win,
                                                        float2 (&v)[FftShape<LOG2M>::Q0][FftShape<LOG2M>::R0]) {
  using S = FftShape<LOG2M>;
  using L0 = SplitPass<S::M, S::R0, 1>;
  constexpr int M = S::M, R0 = S::R0;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
#pragma unroll 1
    for (int n = lane; n < M; n += 64) {
      const int64_t i = pad_index(s0 + 2 * n + half, len, a.pad_mode);
      const float xv = (i >= 0) ? (float)x[i] : 0.0f;
      sf[pidx(n)] = xv * win[2 * n + half];
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < S::Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          if (half) v[q][r].y = sf[L0::rd(j, r)];
          else v[q][r].x = sf[L0::rd(j, r)];
        }
      }
    }
    wave_lds_fence();
  }
}

// FFT of frame `fr` with the last pass in registers: z[q][r] = Z[lane + 64 q
// + BL r]; LDS use: the wave's float buffer of M + M/16 entries at sf. win and
// twp: the window and the per-pass twiddles (global memory or LDS copies).
template <int LOG2M, typename TIn>
__device__ __forceinline__ void frame_fft_split(const TileArgs& a, int clip, int fr, float* sf, int lane,
                                                const float* win, const float2* twp,
                                                float2 (&z)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL],
                                                const float2 (&pf)[FftShape<LOG2M>::Q0][FftShape<LOG2M>::R0],
                                                bool use_pf) {
  using S = FftShape<LOG2M>;
  static_assert(S::REG_SPLIT && S::NP16 > 0, "register-split FFT shapes only");
  constexpr int M = S::M, R0 = S::R0, B0 = S::B0, Q0 = S::Q0;
  using L0 = SplitPass<M, R0, 1>;
  const int64_t len = ldc(a.clip_len, clip);
  const int64_t s0 = (int64_t)fr * a.hop - (a.center ? M : 0);
  lane = opaque(lane);
  float2 v[Q0][R0];
  const TIn* x = static_cast<const TIn*>(a.signal) + ldc(a.clip_off, clip);
  const bool interior = (s0 >= 0) && (s0 + 2 * M <= len);
  if (use_pf) {   // samples prefetched during the previous tile: apply the window
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const float2 w = ldg<float2>(win + 2 * r * B0, (unsigned)j * 8u);
          v[q][r] = make_float2(pf[q][r].x * w.x, pf[q][r].y * w.y);
        }
      }
    }
  } else if (interior && (s0 & 1) == 0) {   // one pair load per complex value
    const TIn* xf = wave_uniform(x + s0);
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const float2 w = ldg<float2>(win + 2 * r * B0, (unsigned)j * 8u);
          const float2 xs = SamplePair<TIn>::load(xf + 2 * r * B0, (unsigned)j * (2u * sizeof(TIn)));
          v[q][r] = make_float2(xs.x * w.x, xs.y * w.y);
        }
      }
    }
  } else if (interior) {   // odd start: element loads
    const TIn* xf = wave_uniform(x + s0);
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int j = lane + 64 * q;
      if (L0::ok(j)) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const float2 w = ldg<float2>(win + 2 * r * B0, (unsigned)j * 8u);
          const unsigned o = (unsigned)j * (2u * sizeof(TIn));
          v[q][r] = make_float2(SamplePair<TIn>::one(xf + 2 * r * B0, o) * w.x,
                                SamplePair<TIn>::one(xf + 2 * r * B0 + 1, o) * w.y);
        }
      }
    }
  } else {
    frame_load_padded_split<LOG2M, TIn>(a, x, len, s0, sf, lane, win, v);
  }
#pragma unroll
  for (int q = 0; q < Q0; ++q) {
    const int j = lane + 64 * q;
    if (L0::ok(j)) Dft<R0>::run(v[q]);
  }
  SplitChain<LOG2M, 1, L0, R0>::run(sf, twp, lane, v, z);
}

// Real-FFT split in registers. Lane l owns Z[k], k = l + 64 q + BL r; its
// mirror Z[M-k] sits in lane (64-l) mod 64 at register (QL-1-q, RL-1-r)
// (lane 0 pairs with itself at (QL-q, RL-1-r), or (0, (RL-r) mod RL) for
// q = 0), fetched with ds_bpermute. Returns 2 X[k] with X[k] = E - i W_N^k O
// as below (the factor 2 saves the halvings and is removed by the caller);
// lane 0 also returns 2 X[M] = 2 (Re Z[0] - Im Z[0]).
template <int LOG2M>
__device__ __forceinline__ void real_split_regs(const float2 (&z)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL],
                                                const float2* __restrict__ twN, int lane_in,
                                                float2 (&x)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL],
                                                float& nyq_re) {
  const int lane = opaque(lane_in);
  using S = FftShape<LOG2M>;
  constexpr int QL = S::QL, RL = S::RL, BL = S::BL;
  const int paddr = ((64 - lane) & 63) << 2;
#pragma unroll
  for (int q = 0; q < QL; ++q) {
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const float2 src = z[QL - 1 - q][RL - 1 - r];
      float2 bz = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(paddr, __float_as_int(src.x))),
                              __int_as_float(__builtin_amdgcn_ds_bpermute(paddr, __float_as_int(src.y))));
      const float2 own = (q == 0) ? z[0][(RL - r) % RL] : z[(QL - q) % QL][RL - 1 - r];
      if (lane == 0) bz = own;
      const float2 A = z[q][r];
      const float2 E = make_float2(A.x + bz.x, A.y - bz.y);   // 2E
      const float2 O = make_float2(A.x - bz.x, A.y + bz.y);   // 2O
      // W_N^k, k = lane + 64 q + BL r
      const float2 T = cmul(O, ldg<float2>(twN + 64 * q + BL * r, (unsigned)lane * 8u));
      x[q][r] = make_float2(E.x + T.y, E.y - T.x);            // 2 X[k]
    }
  }
  nyq_re = 2.0f * (z[0][0].x - z[0][0].y);                   // 2 X[M]
}

// Real-FFT split in registers, one mirror pair per computation. Lane l
// takes the pairs (k, M-k) with k = l + 64 q + BL r for r < RL/2 (each pair
// of the spectrum exactly once: the mirror of such a k sits at r' >= RL/2)
// and returns 2 X[k] in xk[q][r] and 2 X[M-k] = 2 conj(E + i W_N^k O) in
// xm[q][r] (lane 0, q = r = 0: 2 X[0] and the Nyquist 2 X[M]); lane 0 also
// returns xh = 2 X[M/2] = 2 conj(Z[M/2]), the one bin that is its own mirror
// besides DC. Half the mirror fetches, twiddle loads and arithmetic of
// computing every bin on its own lane.
template <int LOG2M>
__device__ __forceinline__ void real_split_pairs(const float2 (&z)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL],
                                                 const float2* __restrict__ twN, int lane_in,
                                                 float2 (&xk)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL / 2],
                                                 float2 (&xm)[FftShape<LOG2M>::QL][FftShape<LOG2M>::RL / 2],
                                                 float2& xh) {
  const int lane = opaque(lane_in);
  using S = FftShape<LOG2M>;
  constexpr int QL = S::QL, RL = S::RL, RH = RL / 2, BL = S::BL;
  static_assert(RL >= 2 && RL % 2 == 0, "pairing needs an even last radix");
  const int paddr = ((64 - lane) & 63) << 2;
#pragma unroll
  for (int q = 0; q < QL; ++q) {
#pragma unroll
    for (int r = 0; r < RH; ++r) {
      const float2 src = z[QL - 1 - q][RL - 1 - r];
      float2 bz = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(paddr, __float_as_int(src.x))),
                              __int_as_float(__builtin_amdgcn_ds_bpermute(paddr, __float_as_int(src.y))));
      const float2 own = (q == 0) ? z[0][(RL - r) % RL] : z[(QL - q) % QL][RL - 1 - r];
      if (lane == 0) bz = own;
      const float2 A = z[q][r];
      const float2 E = make_float2(A.x + bz.x, A.y - bz.y);   // 2E
      const float2 O = make_float2(A.x - bz.x, A.y + bz.y);   // 2O
      // W_N^k, k = lane + 64 q + BL r
      const float2 T = cmul(O, ldg<float2>(twN + 64 * q + BL * r, (unsigned)lane * 8u));
      xk[q][r] = make_float2(E.x + T.y, E.y - T.x);            // 2 X[k]
      xm[q][r] = make_float2(E.x - T.y, -(E.y + T.x));         // 2 X[M-k]
    }
  }
  xh = make_float2(2.0f * z[0][RH].x, -2.0f * z[0][RH].y);    // lane 0: 2 X[M/2]
}

// Real-FFT split: from Z (M complex) to X[k] and X[M-k] for k = lane + 64 i.
// X[k] = E - i W_N^k O,  X[M-k] = conj(E + i W_N^k O),
// E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / 2.
template <int LOG2M>
__device__ __forceinline__ void real_split(const float2* scr, const float2* __restrict__ twN, int k,
                                           float2& xk, float2& xmk) {
  constexpr int M = 1 << LOG2M;
  const float2 A = scr[pidx(k)];
  const float2 Bz = scr[pidx((M - k) & (M - 1))];
  const float2 Bc = make_float2(Bz.x, -Bz.y);
  const float2 E = make_float2(0.5f * (A.x + Bc.x), 0.5f * (A.y + Bc.y));
  const float2 O = make_float2(0.5f * (A.x - Bc.x), 0.5f * (A.y - Bc.y));
  const float2 T = cmul(O, twN[k]);
  xk = make_float2(E.x + T.y, E.y - T.x);
  xmk = make_float2(E.x - T.y, -(E.y + T.x));
}

// |X|^power from |X|^2: power 2 and 1 exactly; otherwise 2^(p/2 log2 |X|^2)
// with the hardware log2/exp2. Callers on the hot path branch on the power
// once per frame and call this only for powers other than 2.
__device__ __forceinline__ float pow_from_p2(float p2, float power) {
  if (power == 2.0f) return p2;
  if (power == 1.0f) return sqrtf(p2);
  return (p2 > 0.0f) ? __builtin_amdgcn_exp2f(0.5f * power * __builtin_amdgcn_logf(p2)) : 0.0f;
}

// 10 log10(v) = 10 log10(2) * log2(v), hardware log2 (v >= amin > 0 here)
__device__ __forceinline__ float db10(float v) { return 3.0102999566398120f * __builtin_amdgcn_logf(v); }

// Maximum / minimum over the 64 lanes of a wave (all lanes active), returned
// in every lane: DPP steps inside each row of 16 lanes (VALU only, no LDS
// round trips), then the four row results through readlane.
template <bool MAX, int CTRL>
__device__ __forceinline__ float dpp_step(float v) {
  const float o = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
  return MAX ? fmaxf(v, o) : fminf(v, o);
}
template <bool MAX>
__device__ __forceinline__ float wave_reduce(float v) {
  v = dpp_step<MAX, 0xB1>(v);    // quad_perm [1, 0, 3, 2]
  v = dpp_step<MAX, 0x4E>(v);    // quad_perm [2, 3, 0, 1]
  v = dpp_step<MAX, 0x141>(v);   // row_half_mirror
  v = dpp_step<MAX, 0x140>(v);   // row_mirror: every lane holds its row's result
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return MAX ? fmaxf(fmaxf(r0, r1), fmaxf(r2, r3)) : fminf(fminf(r0, r1), fminf(r2, r3));
}
// Sum over the 64 lanes, in every lane: the same DPP pattern with adds (each
// step pairs disjoint groups, so every lane ends with the same value).
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
  v = dpp_add<0xB1>(v);
  v = dpp_add<0x4E>(v);
  v = dpp_add<0x141>(v);
  v = dpp_add<0x140>(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return (r0 + r1) + (r2 + r3);
}
__device__ __forceinline__ float wave_max(float v) { return wave_reduce<true>(v); }
__device__ __forceinline__ float wave_min(float v) { return wave_reduce<false>(v); }

// Order-preserving integer image of a float (x < y <=> fkey(x) < fkey(y) for
// non-NaN x, y; -0 sorts just below +0) and its inverse. Integer max / min
// need no NaN canonicalisation, so a DPP step is one v_max_i32_dpp.
__device__ __forceinline__ int fkey(float x) {
  const int i = __float_as_int(x);
  return i ^ ((i >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float fkey_inv(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
template <bool MAX, int CTRL>
__device__ __forceinline__ int dpp_step_i(int v) {
  const int o = __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
  return MAX ? max(v, o) : min(v, o);
}
// Max / min of an int over the 64 lanes of a wave, wave-uniform result.
template <bool MAX>
__device__ __forceinline__ int wave_reduce_i(int v) {
  v = dpp_step_i<MAX, 0xB1>(v);    // quad_perm [1, 0, 3, 2]
  v = dpp_step_i<MAX, 0x4E>(v);    // quad_perm [2, 3, 0, 1]
  v = dpp_step_i<MAX, 0x141>(v);   // row_half_mirror
  v = dpp_step_i<MAX, 0x140>(v);   // row_mirror: every lane holds its row's result
  const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return MAX ? max(max(r0, r1), max(r2, r3)) : min(min(r0, r1), min(r2, r3));
}
// Maximum of non-negative floats (their bit patterns order like the values).
__device__ __forceinline__ float wave_max_nonneg(float v) {
  return __int_as_float(wave_reduce_i<true>(__float_as_int(v)));
}

}  // namespace mm
