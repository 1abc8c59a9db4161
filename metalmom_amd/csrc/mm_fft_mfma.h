// Real FFT of one 2048-sample frame on the matrix cores (16-bit mel / MFCC
// kernels, n_fft 2048).
//
// The frame is packed as M = 1024 complex points z[n] = x[2n] + i x[2n+1]
// (the same packing as the register-split FFT) and transformed in four steps,
// M = 32 x 32 with n = n1 + 32 n2 and k = k2 + 32 k1, each 32-point DFT
// split once by parity (a radix-2 step on the VALU, then two 16-point
// contractions on the matrix cores):
//   S[n1, m] = z[n1 + 32 m] + z[n1 + 32 m + 512], D = their difference  (m < 16)
//   Y[n1, 2q] = sum_m W_16^(q m) S[n1, m],  Y[n1, 2q+1] = sum_m W_32^((2q+1) m) D[n1, m]
//                                                                   (MFMA stage 1)
//   Y' = Y[n1, k2] W_1024^(n1 k2)                                    (VALU twiddle)
//   S2[r, k2] = Y'[r, k2] + Y'[r + 16, k2], D2 = their difference     (r < 16)
//   Z[k2 + 64 u] = sum_r W_16^(u r) S2[r, k2],
//   Z[k2 + 32 (2u+1)] = sum_r W_32^((2u+1) r) D2[r, k2]               (MFMA stage 2)
// Every contraction is a 32 x 32 real matrix product ([re | im] blocks of a
// 16 x 16 complex one) over 32 columns, on v_mfma_f32_16x16x32_f16 with FP32
// accumulation: 24 MFMAs per stage, half of what one 64 x 64 product per
// stage takes. Each FP32 operand x is split as x = hi + lo (two FP16 values,
// ~22 significant bits together) and a product is hi*hi' + lo*hi' + hi*lo'
// (the lo*lo' term, ~2^-20 of the product, is dropped): the transform keeps
// about 2^-20 of the frame's largest value, far below the 16-bit rounding of
// the staged spectrum (bf16: 2^-9, fp16: 2^-11 relative per bin) that
// follows. The frame is scaled by a power of two when its largest sample lies
// outside [2^-4, 2^9), so every FP16 operand stays finite (|S2| <= 2 * 32
// sqrt(2) 2^9 < 65504) and small values keep their precision; the scale is
// returned and removed exactly by the caller.
//
// MFMA operand layout of v_mfma_f32_16x16x32_f16 (lane l, c = l % 16,
// g = l / 16): A[row c][k = 8g + e], B[k = 8g + e][col c], e = 0..7;
// C[row 4g + j][col c], j = 0..3. The 32-wide K dimension holds 16 complex
// points: slot (g, e) is part e / 4 (0 re, 1 im) of point 4 g + (e % 4).
// Stage 1: A = S or D (rows n1 = 16 mt + c), B = the constant [[Wr, Wi],
// [-Wi, Wr]] of its parity p; its outputs (rows n1 = 16 mt + 4 g + j, columns
// k2 = 2 c + p) are, after the twiddle and the second radix-2 step, already
// the stage-2 B operands of the same lane (K = r = 4 g + j, column k2). The
// stage-2 A operand of output parity p1 is the transpose of the stage-1 B of
// parity p1, i.e. the same registers: one 8 KB operand table serves both.
//
// Lane l ends with Z[k2 + 32 k1] for k2 = 2 c + p, k1 = 2 (4 g + j) + p1
// (p, p1 in {0, 1}, j = 0..3). Mirror bins 1024 - k have the same k2 parity,
// so the real split X[k] = E - i W_2048^k O, X[M-k] = conj(E + i W_2048^k O),
// E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2 pairs values of
// one column parity: lane l computes the pairs of its 8 slots s = 4 p + j
// (bin mfft_bin(l, s)) and fetches each mirror partner with ds_bpermute from
// lane mfft_partner(l, s). Lane 0 also computes X[0], X[M] (slot 0; Z[0] is
// its own partner) and lane 32 X[M/2].
#pragma once

#include "mm_fft_dev.h"

namespace mm {

// The table layout (kMf*Bytes) and mfft_kernel are in mm_common.h.

// Whether slot s of lane l computes the pair of its p1 = 1 value (the
// column k2 = 0 pairs k1 with 32 - k1, which has the same parity: lanes 0
// and 16 take its even k1, lanes 32 and 48 its odd k1).
__host__ __device__ constexpr bool mfft_own_odd(int l, int s) { return s < 4 && (l & 15) == 0 && l >= 32; }

// Bin k of slot s = 4 p + j of lane l (its mirror is 1024 - k).
__host__ __device__ inline int mfft_bin(int l, int s) {
  const int c = l & 15, g = l >> 4, p = s >> 2, j = s & 3;
  return (2 * c + p) + 32 * (2 * (4 * g + j) + (mfft_own_odd(l, s) ? 1 : 0));
}

// The lane that holds the mirror of slot s of lane l (l itself for DC).
__host__ __device__ inline int mfft_partner(int l, int s) {
  const int c = l & 15, p = s >> 2, j = s & 3;
  if (p == 1) return 63 - l;
  if (c != 0 || (j == 0 && l < 32)) return (64 - l) & 63;
  return 48 - l;
}

// The value a lane sends in round s (= what its partners need there):
// (p1, j) of its Z(p = s / 4, p1, j).
__host__ __device__ inline void mfft_send(int l, int s, int& p1, int& j) {
  const int p = s >> 2, js = s & 3;
  if (p == 0 && (l & 15) == 0 && l >= 32) {
    p1 = 0;
    j = (4 - js) & 3;
  } else {
    p1 = 1;
    j = 3 - js;
  }
}

// Point index of sample pair i = 4 (2 mt + b) + jj of lane l: z[16 mt + c +
// 32 (16 b + 4 g + jj)] (the radix-2 step adds b = 0 and b = 1 of the same
// lane: points 512 apart).
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
    // operand of parity pp, output part t: element e of lane (c, g) is the
    // entry for input part e / 4 of point m = 4 g + e % 4 and output q = c,
    // W = W_32^((2 q + pp) m)
    r.half = true;
    const int lane = p & 63, t = (p >> 7) & 1, pp = p >> 8;
    const int c = lane & 15, g = lane >> 4;
    for (int e = 0; e < 8; ++e) {
      const int part = e >> 2, m = 4 * g + (e & 3);
      const int q = ((2 * c + pp) * m) & 31;
      const double wr = cos(pi2 * q / 32.0), wi = -sin(pi2 * q / 32.0);
      // [[Wr, Wi], [-Wi, Wr]]: rows (re, im) of the input, columns (re, im) of the output
      r.v[e] = (part == 0) ? (t == 0 ? wr : wi) : (t == 0 ? -wi : wr);
    }
    return r;
  }
  r.half = false;
  if (p < (kMfOpBytes + kMfTw1Bytes) / 16) {
    // W_1024^(n1 k2) of (parity pp, row block mt, j = 2 jp + u): n1 = 16 mt + 4 g + j, k2 = 2 c + pp
    const int q = p - kMfOpBytes / 16, lane = q & 63, t2 = q >> 6;
    const int c = lane & 15, g = lane >> 4, pp = t2 >> 2, mt = (t2 >> 1) & 1, jp = t2 & 1;
    for (int u = 0; u < 2; ++u) {
      const int n1 = 16 * mt + 4 * g + 2 * jp + u, k2 = 2 * c + pp;
      const int e = (n1 * k2) & 1023;
      r.v[2 * u] = cos(pi2 * e / 1024.0);
      r.v[2 * u + 1] = -sin(pi2 * e / 1024.0);
    }
    for (int e = 4; e < 8; ++e) r.v[e] = 0.0;
    return r;
  }
  // W_2048^k of the bins of slots 4 pp + 2 jp + u
  const int q = p - (kMfOpBytes + kMfTw1Bytes) / 16, lane = q & 63, s2 = q >> 6;
  for (int u = 0; u < 2; ++u) {
    const int k = mfft_bin(lane, 2 * s2 + u);
    r.v[2 * u] = cos(pi2 * k / 2048.0);
    r.v[2 * u + 1] = -sin(pi2 * k / 2048.0);
  }
  for (int e = 4; e < 8; ++e) r.v[e] = 0.0;
  return r;
}

// Writes 16-byte piece p (0 .. kMfTabBytes / 16 - 1) of the table at tab:
// window pairs from the natural-order window win[n_fft], then the pieces of
// mfft_table_piece (operand pieces: FP16 hi of the double entry, or lo = its
// remainder, by the piece's hi / lo bit).
__host__ __device__ inline void mfft_table_store(int p, const float* win, unsigned char* tab) {
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

// Four complex values (K slots of one lane: re parts then im parts) -> FP16
// hi and lo operands.
__device__ __forceinline__ void mf_pack4(const float2 (&w)[4], float m1, f16x8& hi, f16x8& lo) {
  float x[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    x[j] = w[j].x;
    x[4 + j] = w[j].y;
  }
  mf_split8(x, m1, hi, lo);
}

__device__ __forceinline__ f32x4 mf_mma(const f16x8& a, const f16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// hi*hi' + lo*hi' + hi*lo' into a fresh accumulator
__device__ __forceinline__ f32x4 mf_mma3(const f16x8& ah, const f16x8& al, const f16x8& bh, const f16x8& bl) {
  f32x4 acc = mf_mma(ah, bh, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
  acc = mf_mma(al, bh, acc);
  return mf_mma(ah, bl, acc);
}

// 16 wait states after the MFMAs that wrote the accumulators before any VALU
// instruction reads them (the compiler's own count for
// v_mfma_f32_16x16x32_f16 results is too short; see mfma_settle).
__device__ __forceinline__ void mf_settle4(f32x4& a0, f32x4& a1, f32x4& a2, f32x4& a3) {
  asm volatile("s_nop 7\n\ts_nop 7" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3));
}

// Power-of-two exponent es that brings the frame's largest |sample| m into
// [2^8, 2^9), or 0 when m is in [2^-4, 2^9) already (or m is 0, Inf, NaN).
__device__ __forceinline__ int mf_scale_exp(float m) {
  if (m >= 0.0625f && m < 512.0f) return 0;
  if (!(m > 0.0f) || m > 3.0e38f) return 0;
  return 9 - __builtin_amdgcn_frexp_expf(m);
}

// The transform. v[i]: windowed pair z[mfft_point(lane, i)] (already scaled
// by 2^es); tab: the table (LDS). Returns 2 X[k] in xk[s], 2 X[1024-k] in
// xm[s] for k = mfft_bin(lane, s) (lane 0, s = 0: 2 X[0] and 2 X[1024]) and,
// in lane 32, xh = 2 X[512].
//
// Order: each group of 12 MFMAs is issued next to VALU work it does not
// depend on (the next operand split, the previous parity's twiddle, the real
// split of the finished column parity), so that instructions of one wave
// overlap the matrix pipeline; a wave's MFMAs do not overlap other waves'
// VALU instructions.
__device__ __forceinline__ void mfft2048(const float2 (&v)[16], const unsigned char* tab, int lane_in,
                                         float2 (&xk)[8], float2 (&xm)[8], float2& xh) {
  const int lane = opaque(lane_in);
  const float m1 = mf_neg_one();
  const uint4* op = reinterpret_cast<const uint4*>(tab + kMfOpOff);
  // constant operand of parity p, part t (hi: h = 0, lo: h = 1)
  auto opnd = [&](int p, int t, int h) { return __builtin_bit_cast(f16x8, op[((p * 2 + t) * 2 + h) * 64 + lane]); };
  const float4* tw1 = reinterpret_cast<const float4*>(tab + kMfTw1Off);
  const float4* tw2 = reinterpret_cast<const float4*>(tab + kMfTw2Off);

  // ---- stage-1 radix-2 step: S and D of rows 16 mt + c ----
  float2 sd[2][2][4];   // [p: S, D][mt][jj]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const float2 a = v[4 * (2 * mt) + jj], b = v[4 * (2 * mt + 1) + jj];
      sd[0][mt][jj] = make_float2(a.x + b.x, a.y + b.y);
      sd[1][mt][jj] = make_float2(a.x - b.x, a.y - b.y);
    }
  f16x8 ah[2][2], al[2][2];   // [p][mt]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) mf_pack4(sd[0][mt], m1, ah[0][mt], al[0][mt]);

  // ---- stage 1: acc[p][mt][t] = part t of Y[16 mt + 4 g + j][2 c + p] ----
  f32x4 acc[2][2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const f16x8 bh0 = opnd(p, 0, 0), bl0 = opnd(p, 0, 1), bh1 = opnd(p, 1, 0), bl1 = opnd(p, 1, 1);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      acc[p][mt][0] = mf_mma3(ah[p][mt], al[p][mt], bh0, bl0);
      acc[p][mt][1] = mf_mma3(ah[p][mt], al[p][mt], bh1, bl1);
    }
    // next parity's operands, while these MFMAs run
    if (p == 0) {
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) mf_pack4(sd[1][mt], m1, ah[1][mt], al[1][mt]);
    }
  }

  // ---- twiddle + stage-2 radix-2 step of column parity p, then stage 2 ----
  // B operands [p][s] (s = 0: S2, feeds the even outputs k1 = 2u; s = 1: D2,
  // the odd ones); acc2[p][p1][t1] = part t1 of Z[(2 c + p) + 32 (2 (4 g + j) + p1)]
  f16x8 ch[2][2], cl[2][2];
  f32x4 acc2[2][2][2];
  auto twiddle = [&](int p) {
    float2 y[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int jp = 0; jp < 2; ++jp) {
        const float4 w = tw1[((p * 2 + mt) * 2 + jp) * 64 + lane];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int j = 2 * jp + u;
          const float wr = u ? w.z : w.x, wi = u ? w.w : w.y;
          const float yr = acc[p][mt][0][j], yi = acc[p][mt][1][j];
          y[mt][j] = make_float2(__builtin_fmaf(yr, wr, -yi * wi), __builtin_fmaf(yr, wi, yi * wr));
        }
      }
    float2 s2[4], d2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      s2[j] = make_float2(y[0][j].x + y[1][j].x, y[0][j].y + y[1][j].y);
      d2[j] = make_float2(y[0][j].x - y[1][j].x, y[0][j].y - y[1][j].y);
    }
    mf_pack4(s2, m1, ch[p][0], cl[p][0]);
    mf_pack4(d2, m1, ch[p][1], cl[p][1]);
  };
  auto stage2 = [&](int p) {
#pragma unroll
    for (int p1 = 0; p1 < 2; ++p1)
#pragma unroll
      for (int t1 = 0; t1 < 2; ++t1)
        acc2[p][p1][t1] = mf_mma3(opnd(p1, t1, 0), opnd(p1, t1, 1), ch[p][p1], cl[p][p1]);
  };
  // ---- paired real split of column parity p ----
  const bool c0odd = mfft_own_odd(lane, 0);   // lanes 32, 48 (same for every p = 0 slot)
  const int base = (64 - lane) << 2;          // partner of the regular p = 0 slots (bpermute byte address)
  const int pa0x = ((lane & 15) == 0) ? base - 64 : base;             // p = 0, j > 0
  const int pa00 = ((lane & 15) == 0 && lane >= 32) ? base - 64 : base;   // p = 0, j = 0
  const int pa1 = (63 - lane) << 2;
  auto split = [&](int p) {
    auto Zv = [&](int p1, int j) { return make_float2(acc2[p][p1][0][j], acc2[p][p1][1][j]); };
#pragma unroll
    for (int jp = 0; jp < 2; ++jp) {
      const float4 w = tw2[(p * 2 + jp) * 64 + lane];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int j = 2 * jp + u, s = 4 * p + j;
        float2 own = Zv(0, j), snd = Zv(1, 3 - j);
        int addr = pa1;
        if (p == 0) {
          own = c0odd ? Zv(1, j) : own;
          snd = c0odd ? Zv(0, (4 - j) & 3) : snd;
          addr = (j == 0) ? pa00 : pa0x;
        }
        float2 B = make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(snd.x))),
                               __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(snd.y))));
        if (s == 0) B = (lane == 0) ? own : B;   // DC: Z[0] is its own mirror
        const float2 E = make_float2(own.x + B.x, own.y - B.y);   // 2E
        const float2 O = make_float2(own.x - B.x, own.y + B.y);   // 2O
        const float wr = u ? w.z : w.x, wi = u ? w.w : w.y;
        const float2 T = make_float2(__builtin_fmaf(O.x, wr, -O.y * wi), __builtin_fmaf(O.x, wi, O.y * wr));
        xk[s] = make_float2(E.x + T.y, E.y - T.x);          // 2 X[k]
        xm[s] = make_float2(E.x - T.y, -(E.y + T.x));       // 2 X[M-k]
      }
    }
  };

  mf_settle4(acc[0][0][0], acc[0][0][1], acc[0][1][0], acc[0][1][1]);
  twiddle(0);
  stage2(0);
  mf_settle4(acc[1][0][0], acc[1][0][1], acc[1][1][0], acc[1][1][1]);
  twiddle(1);
  stage2(1);
  mf_settle4(acc2[0][0][0], acc2[0][0][1], acc2[0][1][0], acc2[0][1][1]);
  split(0);
  mf_settle4(acc2[1][0][0], acc2[1][0][1], acc2[1][1][0], acc2[1][1][1]);
  split(1);
  // X[512] = conj(Z[512]); Z[512] is Z(p = 0, p1 = 0, j = 0) of lane 32
  xh = make_float2(2.0f * acc2[0][0][0][0], -2.0f * acc2[0][0][1][0]);
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
