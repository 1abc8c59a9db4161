// Real FFT of one 2048-sample frame on the matrix cores (16-bit mel / MFCC
// kernels, n_fft 2048).
//
// The frame is packed as M = 1024 complex points z[n] = x[2n] + i x[2n+1]
// (the same packing as the register-split FFT) and transformed in four steps,
// M = 32 x 32 with n = n1 + 32 n2 and k = k2 + 32 k1:
//   Y[n1, k2] = sum_n2 z[n1 + 32 n2] W_32^(n2 k2)          (MFMA stage 1)
//   Y'        = Y[n1, k2] W_1024^(n1 k2)                    (VALU twiddle)
//   Z[k2 + 32 k1] = sum_n1 W_32^(n1 k1) Y'[n1, k2]           (MFMA stage 2)
// Both DFT stages are 64 x 64 real matrix products ([re | im] blocks) on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation. Each FP32 operand x is
// split as x = hi + lo (two FP16 values, ~22 significant bits together) and a
// product is hi*hi' + lo*hi' + hi*lo' (the lo*lo' term, ~2^-20 of the
// product, is dropped): the transform keeps about 2^-20 of the frame's
// largest value, far below the 16-bit rounding of the staged spectrum
// (bf16: 2^-9, fp16: 2^-11 relative per bin) that follows. The frame is
// scaled by a power of two when its largest sample lies outside [2^-4, 2^10)
// so that every FP16 intermediate stays finite (|Y| <= 32 sqrt(2) 2^10) and
// small values keep their precision; the scale is returned and removed
// exactly by the caller.
//
// MFMA operand layout of v_mfma_f32_16x16x32_f16 (lane l, c = l % 16,
// g = l / 16): A[row c][k = 8g + e], B[k = 8g + e][col c], e = 0..7;
// C[row 4g + j][col c], j = 0..3. Inside a 64-wide K dimension (two blocks of
// 32) the K slot (block b, group g, element e) stands for part e / 4 (0 re,
// 1 im) of point 16 b + 4 g + (e % 4). With that order the stage-1 output
// tiles are already the stage-2 B operands of the same lane (no exchange),
// and the constant operand is the same registers in both stages: stage 1's B
// is [[Wr, Wi], [-Wi, Wr]] and stage 2's A its transpose, and W_32 is
// symmetric.
//
// Lane l ends with Z[k2 + 32 k1] for k1 = 16 h1 + 4 g + j, k2 = 16 h2 + c
// (h1, h2 in {0, 1}, j = 0..3). The real split X[k] = E - i W_2048^k O,
// X[M-k] = conj(E + i W_2048^k O), E = (Z[k] + conj Z[M-k]) / 2,
// O = (Z[k] - conj Z[M-k]) / 2 takes each mirror pair (k, M-k) once: lane l
// computes 8 pairs (slots s = 4 h1 + j, bin mfft_bin(l, s)) and fetches the
// mirror partner with ds_bpermute. Lane 0 also computes X[0], X[M] (slot 4)
// and X[M/2].
#pragma once

#include "mm_fft_dev.h"

namespace mm {

// The table layout (kMf*Bytes) and mfft_kernel are in mm_common.h.

// Bin k of slot s = 4 h1 + j of lane l (its mirror is 1024 - k).
__host__ __device__ inline int mfft_bin(int l, int s) {
  const int c = l & 15, g = l >> 4, h1 = s >> 2, j = s & 3;
  if (c != 0) return c + 32 * (16 * h1 + 4 * g + j);
  if (h1 == 0) return 16 + 32 * (4 * g + j);   // column k2 = 16
  if (l == 0 && j == 0) return 0;              // DC (with the Nyquist bin as its mirror)
  return 32 * (16 + 4 * g + j);                // column k2 = 0
}

// Point index of sample pair i = 4 (2 mt + b) + jj of lane l: the stage-1 A
// operand of tile row block mt, K block b holds z[16 mt + c + 32 (16 b + 4 g + jj)].
__host__ __device__ inline int mfft_point(int l, int i) {
  const int c = l & 15, g = l >> 4, mt = i >> 3, b = (i >> 2) & 1, jj = i & 3;
  return 16 * mt + c + 32 * (16 * b + 4 * g + jj);
}

// Entry of 16-byte piece p of the table after its window block (p = 0 ..
// (kMfTabBytes - kMfWinBytes) / 16 - 1), in double precision: 8 FP16 values
// (operand pieces) or 2 float2 (twiddles).
struct MfPiece {
  bool half;
  double v[8];
};
__host__ __device__ inline MfPiece mfft_table_piece(int p) {
  MfPiece r;
  const double pi2 = 6.283185307179586476925286766559;
  if (p < kMfOpBytes / 16) {
    r.half = true;
    const int lane = p & 63, b = (p >> 7) & 1, t = p >> 8;
    const int c = lane & 15, g = lane >> 4;
    const int col = 16 * t + c, outp = col >> 5, k2 = col & 31;
    for (int e = 0; e < 8; ++e) {
      const int part = e >> 2, n2 = 16 * b + 4 * g + (e & 3);
      const int q = (n2 * k2) & 31;
      const double wr = cos(pi2 * q / 32.0), wi = -sin(pi2 * q / 32.0);
      // [[Wr, Wi], [-Wi, Wr]]: rows (re, im) of the input, columns (re, im) of the output
      r.v[e] = (part == 0) ? (outp == 0 ? wr : wi) : (outp == 0 ? -wi : wr);
    }
    return r;
  }
  r.half = false;
  if (p < (kMfOpBytes + kMfTw1Bytes) / 16) {
    const int q = p - kMfOpBytes / 16, lane = q & 63, t2 = q >> 6;
    const int c = lane & 15, g = lane >> 4;
    for (int u = 0; u < 2; ++u) {
      const int t = 2 * t2 + u, mt = t >> 3, kh = (t >> 2) & 1, j = t & 3;
      const int n1 = 16 * mt + 4 * g + j, k2 = 16 * kh + c;
      const int e = (n1 * k2) & 1023;
      r.v[2 * u] = cos(pi2 * e / 1024.0);
      r.v[2 * u + 1] = -sin(pi2 * e / 1024.0);
    }
    return r;
  }
  const int q = p - (kMfOpBytes + kMfTw1Bytes) / 16, lane = q & 63, s2 = q >> 6;
  for (int u = 0; u < 2; ++u) {
    const int k = mfft_bin(lane, 2 * s2 + u);
    r.v[2 * u] = cos(pi2 * k / 2048.0);
    r.v[2 * u + 1] = -sin(pi2 * k / 2048.0);
  }
  for (int e = 4; e < 8; ++e) r.v[e] = 0.0;
  return r;
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// -1.0f in a scalar register the compiler cannot fold: the low-part
// subtraction x - (float)hi then becomes one v_fma_mix_f32 (FP16 operand
// widened inside the FMA) instead of a conversion plus a subtraction.
__device__ __forceinline__ float mf_neg_one() {
  float m;
  __asm__ __volatile__("s_mov_b32 %0, 0xbf800000" : "=s"(m));
  return m;
}

// x[0..7] -> FP16 hi and lo operands (element e = x[e]); conversions round
// toward zero (the low part is exact in FP32: x - hi has at most 14 bits).
__device__ __forceinline__ void mf_split8(const float (&x)[8], float m1, f16x8& hi, f16x8& lo) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const h2v h = __builtin_bit_cast(h2v, __builtin_amdgcn_cvt_pkrtz(x[2 * p], x[2 * p + 1]));
    const float l0 = __builtin_fmaf((float)h.x, m1, x[2 * p]);
    const float l1 = __builtin_fmaf((float)h.y, m1, x[2 * p + 1]);
    const h2v l = __builtin_bit_cast(h2v, __builtin_amdgcn_cvt_pkrtz(l0, l1));
    hi[2 * p] = h.x;
    hi[2 * p + 1] = h.y;
    lo[2 * p] = l.x;
    lo[2 * p + 1] = l.y;
  }
}

// Writes 16-byte piece p (0 .. kMfTabBytes / 16 - 1) of the table at tab:
// window pairs from the natural-order window win[n_fft], then the pieces of
// mfft_table_piece (operand pieces: FP16 hi of the double entry, or lo = its
// remainder, by the piece's hi / lo bit).
__device__ inline void mfft_table_store(int p, const float* win, unsigned char* tab) {
  if (p < kMfWinBytes / 16) {
    const int lane = p & 63, pp = p >> 6;
    float f[4];
    for (int u = 0; u < 2; ++u) {
      const int n = mfft_point(lane, 2 * pp + u);
      f[2 * u] = win[2 * n];
      f[2 * u + 1] = win[2 * n + 1];
    }
    __builtin_memcpy(tab + 16 * p, f, 16);
    return;
  }
  tab += kMfWinBytes;
  p -= kMfWinBytes / 16;
  const MfPiece r = mfft_table_piece(p);
  if (r.half) {
    const bool lo = (p >> 6) & 1;
    _Float16 h[8];
    for (int e = 0; e < 8; ++e) {
      const _Float16 hi = (_Float16)r.v[e];
      h[e] = lo ? (_Float16)(r.v[e] - (double)hi) : hi;
    }
    __builtin_memcpy(tab + 16 * p, h, 16);
  } else {
    float f[4];
    for (int e = 0; e < 4; ++e) f[e] = (float)r.v[e];
    __builtin_memcpy(tab + 16 * p, f, 16);
  }
}

__device__ __forceinline__ f32x4 mf_mma(const f16x8& a, const f16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// 16 wait states after the MFMAs that wrote the accumulators before any VALU
// instruction reads them (the compiler's own count for
// v_mfma_f32_16x16x32_f16 results is too short; see mfma_settle).
__device__ __forceinline__ void mf_settle8(f32x4& a0, f32x4& a1, f32x4& a2, f32x4& a3, f32x4& a4, f32x4& a5,
                                           f32x4& a6, f32x4& a7) {
  asm volatile("s_nop 7\n\ts_nop 7" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7));
}

// Power-of-two exponent es that brings the frame's largest |sample| m into
// [2^9, 2^10), or 0 when m is in [2^-4, 2^10) already (or m is 0, Inf, NaN).
__device__ __forceinline__ int mf_scale_exp(float m) {
  if (m >= 0.0625f && m < 1024.0f) return 0;
  if (!(m > 0.0f) || m > 3.0e38f) return 0;
  return 10 - __builtin_amdgcn_frexp_expf(m);
}

// The transform. v[i]: windowed pair z[mfft_point(lane, i)] (already scaled
// by 2^es); tab: the table (LDS). Returns 2 X[k] in xk[s], 2 X[1024-k] in
// xm[s] for k = mfft_bin(lane, s) (lane 0, s = 4: 2 X[0] and 2 X[1024]) and,
// in lane 0, xh = 2 X[512].
__device__ __forceinline__ void mfft2048(const float2 (&v)[16], const unsigned char* tab, int lane_in,
                                         float2 (&xk)[8], float2 (&xm)[8], float2& xh) {
  const int lane = opaque(lane_in);
  const float m1 = mf_neg_one();
  const uint4* op = reinterpret_cast<const uint4*>(tab + kMfOpOff);
  // ---- stage 1: A = samples [mt][b], B = constant tile t ----
  f16x8 ah[2][2], al[2][2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float x[8];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        x[jj] = v[4 * (2 * mt + b) + jj].x;
        x[4 + jj] = v[4 * (2 * mt + b) + jj].y;
      }
      mf_split8(x, m1, ah[mt][b], al[mt][b]);
    }
  f32x4 acc[2][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f16x8 bh[2], bl[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      bh[b] = __builtin_bit_cast(f16x8, op[((t * 2 + b) * 2 + 0) * 64 + lane]);
      bl[b] = __builtin_bit_cast(f16x8, op[((t * 2 + b) * 2 + 1) * 64 + lane]);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) acc[mt][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) acc[mt][t] = mf_mma(ah[mt][b], bh[b], acc[mt][t]);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) acc[mt][t] = mf_mma(al[mt][b], bh[b], acc[mt][t]);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) acc[mt][t] = mf_mma(ah[mt][b], bl[b], acc[mt][t]);
    }
  }
  mf_settle8(acc[0][0], acc[0][1], acc[0][2], acc[0][3], acc[1][0], acc[1][1], acc[1][2], acc[1][3]);
  // ---- twiddle: lane holds Y[n1 = 16 mt + 4 g + j][k2 = 16 kh + c] (re: tile kh, im: tile kh + 2) ----
  const float4* tw1 = reinterpret_cast<const float4*>(tab + kMfTw1Off);
  f16x8 ch[2][2], cl[2][2];   // stage-2 B operands [kh][block = mt]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {
      float x[8];
#pragma unroll
      for (int jp = 0; jp < 2; ++jp) {
        const float4 w = tw1[((mt * 2 + kh) * 2 + jp) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int j = 2 * jp + u;
          const float wr = u ? w.z : w.x, wi = u ? w.w : w.y;
          const float yr = acc[mt][kh][j], yi = acc[mt][kh + 2][j];
          x[j] = __builtin_fmaf(yr, wr, -yi * wi);
          x[4 + j] = __builtin_fmaf(yr, wi, yi * wr);
        }
      }
      mf_split8(x, m1, ch[kh][mt], cl[kh][mt]);
    }
  // ---- stage 2: A = constant tile t (rows: re / im of k1), B = twiddled Y ----
  f32x4 acc2[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f16x8 wh[2], wl[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      wh[b] = __builtin_bit_cast(f16x8, op[((t * 2 + b) * 2 + 0) * 64 + lane]);
      wl[b] = __builtin_bit_cast(f16x8, op[((t * 2 + b) * 2 + 1) * 64 + lane]);
    }
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) acc2[t][kh] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int kh = 0; kh < 2; ++kh) acc2[t][kh] = mf_mma(wh[b], ch[kh][b], acc2[t][kh]);
#pragma unroll
      for (int kh = 0; kh < 2; ++kh) acc2[t][kh] = mf_mma(wl[b], ch[kh][b], acc2[t][kh]);
#pragma unroll
      for (int kh = 0; kh < 2; ++kh) acc2[t][kh] = mf_mma(wh[b], cl[kh][b], acc2[t][kh]);
    }
  }
  mf_settle8(acc2[0][0], acc2[0][1], acc2[1][0], acc2[1][1], acc2[2][0], acc2[2][1], acc2[3][0], acc2[3][1]);
  // Z(h1, j, h2) = Z[(16 h2 + c) + 32 (16 h1 + 4 g + j)]
  auto Zv = [&](int h1, int j, int h2) { return make_float2(acc2[h1][h2][j], acc2[2 + h1][h2][j]); };
  // ---- paired real split ----
  const int c = lane & 15, g = lane >> 4;
  const bool c0 = (c == 0);
  const int pa = (16 * (3 - g) + ((16 - c) & 15)) << 2;       // mirror lane of every slot
  const int pa10 = c0 ? ((16 * ((4 - g) & 3)) << 2) : pa;     // slot 4 of the column k2 = 0
  const float4* tw2 = reinterpret_cast<const float4*>(tab + kMfTw2Off);
#pragma unroll
  for (int sp = 0; sp < 4; ++sp) {
    const float4 w = tw2[sp * 64 + lane];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int s = 2 * sp + u, h1 = s >> 2, j = s & 3;
      // what this lane sends for slot s, and its own value of the pair
      float2 snd = Zv(1 - h1, 3 - j, 1);
      float2 own = Zv(h1, j, 0);
      if (h1 == 1) {
        const float2 z0 = Zv(0, (4 - j) & 3, 0);   // column 0: k1 <-> 32 - k1
        snd = c0 ? z0 : snd;
        if (j == 0) own = (lane == 0) ? Zv(0, 0, 0) : own;   // lane 0: DC with itself
      } else {
        const float2 z16 = Zv(0, j, 1);   // column 16: k1 <-> 31 - k1
        own = c0 ? z16 : own;
      }
      const int addr = (s == 4) ? pa10 : pa;
      const float2 B = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(snd.x))),
                                   __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(snd.y))));
      const float2 E = make_float2(own.x + B.x, own.y - B.y);   // 2E
      const float2 O = make_float2(own.x - B.x, own.y + B.y);   // 2O
      const float wr = u ? w.z : w.x, wi = u ? w.w : w.y;
      const float2 T = make_float2(__builtin_fmaf(O.x, wr, -O.y * wi), __builtin_fmaf(O.x, wi, O.y * wr));
      xk[s] = make_float2(E.x + T.y, E.y - T.x);          // 2 X[k]
      xm[s] = make_float2(E.x - T.y, -(E.y + T.x));       // 2 X[M-k]
    }
  }
  const float2 z512 = Zv(1, 0, 0);   // lane 0: Z[512]
  xh = make_float2(2.0f * z512.x, -2.0f * z512.y);
}

// Constant part of the offset (in sample pairs) of pair i of every lane:
// mfft_point(l, i) = ((l % 16) + 128 (l / 16)) + mfft_pair_off(i).
__host__ __device__ constexpr int mfft_pair_off(int i) {
  return 16 * (i >> 3) + 32 * (16 * ((i >> 2) & 1) + (i & 3));
}

// Windowed input pairs v[i] = (x[s0 + 2n] w[2n], x[s0 + 2n + 1] w[2n + 1]),
// n = mfft_point(lane, i), of frame fr of a clip: from the prefetched pairs
// pf (use_pf), straight from global memory for interior frames, or through
// pad_index for frames that touch the centre padding. tab: the table (LDS
// copy; its window block), a.window: the natural-order window.
template <typename TIn>
__device__ __forceinline__ void mfft_load(const TileArgs& a, int clip, int fr, int lane, const unsigned char* tab,
                                          float2 (&v)[16], const float2 (&pf)[16], bool use_pf) {
  const int64_t len = ldc(a.clip_len, clip);
  const int64_t s0 = (int64_t)fr * a.hop - (a.center ? 1024 : 0);
  const TIn* x = static_cast<const TIn*>(a.signal) + ldc(a.clip_off, clip);
  const bool interior = (s0 >= 0) && (s0 + 2048 <= len);
  const unsigned lp = (unsigned)((lane & 15) + 128 * (lane >> 4));
  // this lane's window pairs 2p, 2p+1: 16 bytes, lanes contiguous
  auto wpair = [&](int p) { return ldg<float4>(tab + p * 1024, (unsigned)lane * 16u); };
  if (use_pf) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float4 w = wpair(p);
      v[2 * p] = make_float2(pf[2 * p].x * w.x, pf[2 * p].y * w.y);
      v[2 * p + 1] = make_float2(pf[2 * p + 1].x * w.z, pf[2 * p + 1].y * w.w);
    }
  } else if (interior && (s0 & 1) == 0) {
    const TIn* xf = wave_uniform(x + s0);
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float4 w = wpair(p);
      const float2 x0 = SamplePair<TIn>::load(xf + 2 * mfft_pair_off(2 * p), lp * (2u * sizeof(TIn)));
      const float2 x1 = SamplePair<TIn>::load(xf + 2 * mfft_pair_off(2 * p + 1), lp * (2u * sizeof(TIn)));
      v[2 * p] = make_float2(x0.x * w.x, x0.y * w.y);
      v[2 * p + 1] = make_float2(x1.x * w.z, x1.y * w.w);
    }
  } else if (interior) {
    const TIn* xf = wave_uniform(x + s0);
    const unsigned o = lp * (2u * sizeof(TIn));
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float4 w = wpair(p);
      const int c0 = 2 * mfft_pair_off(2 * p), c1 = 2 * mfft_pair_off(2 * p + 1);
      v[2 * p] = make_float2(SamplePair<TIn>::one(xf + c0, o) * w.x, SamplePair<TIn>::one(xf + c0 + 1, o) * w.y);
      v[2 * p + 1] = make_float2(SamplePair<TIn>::one(xf + c1, o) * w.z, SamplePair<TIn>::one(xf + c1 + 1, o) * w.w);
    }
  } else {
    // unrolled: an indexed register array in a rolled loop would be placed
    // in scratch memory for the whole kernel
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int n = (int)lp + mfft_pair_off(i);
      const int64_t i0 = pad_index(s0 + 2 * n, len, a.pad_mode);
      const int64_t i1 = pad_index(s0 + 2 * n + 1, len, a.pad_mode);
      const float x0 = (i0 >= 0) ? (float)x[i0] : 0.0f, x1 = (i1 >= 0) ? (float)x[i1] : 0.0f;
      const float4 w = wpair(i >> 1);
      v[i] = (i & 1) ? make_float2(x0 * w.z, x1 * w.w) : make_float2(x0 * w.x, x1 * w.y);
    }
  }
}

// m = max(m, |a|, |b|) in one instruction (fmaxf over fabsf compiles to
// three per pair: in IEEE mode each operand is first quieted by a max with
// itself). A NaN operand may leave m anything; mf_scale_exp then picks no
// scaling or some power of two, either of which the caller undoes exactly.
__device__ __forceinline__ void mf_absmax2(float& m, float a, float b) {
  __asm__("v_max3_f32 %0, |%1|, |%2|, %0" : "+v"(m) : "v"(a), "v"(b));
}

// The whole frame: windowed pairs -> power-of-two scale -> transform.
// Returns the scale exponent es (the spectrum is 2^es times the frame's).
__device__ __forceinline__ int mfft_frame(float2 (&v)[16], const unsigned char* tab, int lane, float2 (&xk)[8],
                                          float2 (&xm)[8], float2& xh) {
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) mf_absmax2(m, v[i].x, v[i].y);
  const int es = mf_scale_exp(wave_max_nonneg(m));
  if (es != 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = make_float2(ldexpf(v[i].x, es), ldexpf(v[i].y, es));
  }
  mfft2048(v, tab, lane, xk, xm, xh);
  return es;
}

#endif

}  // namespace mm
