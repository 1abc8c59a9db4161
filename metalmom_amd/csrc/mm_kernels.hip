// HIP kernels for gfx950: device-side FP64 table generation, the fused
// frame-tile kernel (framing -> Hann -> real FFT -> |X|^p -> banded mel on
// MFMA -> dB -> DCT-II on MFMA), generic fallbacks for FFT sizes outside the
// fused range, and the standalone dB conversion.
//
// Reference semantics followed (zakkeown/MetalMom):
//   framing / centre padding / frame count  Sources/MetalMomCore/Spectral/STFT.swift:240-254
//   periodic Hann + centred pad/truncate     Filters/Windows.swift:14-48, STFT.swift:44-57
//   real FFT, DC/Nyquist, power, magnitude   STFT.swift:287-329, 140-197
//   Slaney mel filterbank                    Filters/FilterBank.swift:23-73, Convert/Units.swift:55-85
//   FB @ |X|^power                           Spectral/MelSpectrogram.swift:28-142
//   power_to_db with a whole-clip top_db     Spectral/Scaling.swift:73-118
//   DCT-II ortho basis product               Features/MFCC.swift:79-106
#include "mm_common.h"
#include "mm_fft_dev.h"
#include "mm_fft_mfma.h"
#include "mm_kernels.h"

#include <algorithm>

// Waves per SIMD the fused kernel is compiled for: 8-wave workgroups up to
// n_fft 2048 run two per CU (4 waves per SIMD), so their registers are
// capped at 128.
template <int LOG2M, int NW>
constexpr int tile_waves_per_eu() { return (NW == 8 && LOG2M <= 10) ? 4 : 1; }

// Optional phase timing of the fused kernel (build with -DMM_PHASE_TIMING):
// lane 0 of every wave accumulates shader-clock cycles spent in each tile
// phase and stores the sums to TileArgs::phase_cycles at exit. Slots: FFT +
// power, staging, first barrier, mel + dB, extrema reduction, DCT, next-tile
// prefetch, second barrier, cross-block sum.
constexpr int kPhaseSlots = 9;
#ifdef MM_PHASE_TIMING
#define MM_PT_DECL long long pt_acc[kPhaseSlots] = {}; long long pt_last = __builtin_amdgcn_s_memtime();
#define MM_PT(i) do { const long long now_ = __builtin_amdgcn_s_memtime(); pt_acc[i] += now_ - pt_last; pt_last = now_; } while (0)
#define MM_PT_STORE() do { if (a.phase_cycles && lane0 == 0) for (int i_ = 0; i_ < kPhaseSlots; ++i_) a.phase_cycles[(blockIdx.x * NW + wave) * kPhaseSlots + i_] += pt_acc[i_]; } while (0)
#else
#define MM_PT_DECL
#define MM_PT(i) do { } while (0)
#define MM_PT_STORE() do { } while (0)
#endif

namespace mm {

// ---------------------------------------------------------------------------
// the fused tile kernel
// ---------------------------------------------------------------------------
//
// One workgroup = NW wavefronts; one tile = NW*FPW consecutive frames of one
// clip (16 for the mel/MFCC outputs: the MFMA N dimension). Wave w
// transforms frames w*FPW .. w*FPW+FPW-1 of the tile in its own LDS region,
// keeping each spectrum in registers until the FFT scratch under it is free,
// then writes it into its staging row(s) in the same region. After a
// workgroup barrier:
//   * STFT outputs are written transposed ([bins, frames], TF-frame runs);
//   * mel: each wave takes the blocks of 8 mels that mel_block_assign gave
//     it (longest first, balanced by bin chunks) and contracts the banded
//     filterbank with the staged spectra on MFMA
//     (mel^T[8 mels x 16 frames] = FB[8 x K] * P^T[K x 16], MFMA rows 8..15
//     repeating rows 0..7);
//   * MFCC: dB in registers, then DCT-II on MFMA with the mel registers fed
//     straight in as the B operand (the DCT table is stored in the matching
//     K order), partial sums reduced across waves in LDS in fixed order.
// Persistent: workgroup g processes tiles [g*C, (g+1)*C).
//
// PIPE (SHAPE_W16P, 16-bit MFCC with the matrix-core FFT): the staging rows,
// frame exponents, per-wave extrema and DCT partials exist twice and
// processed tiles alternate between the two sets, so one barrier per tile
// suffices: after it a wave first finishes the previous tile (top_db
// bookkeeping and block sum, waves 0-3) and contracts this one, then goes on
// to the next FFT while other waves are still contracting.
template <int LOG2M, int MODE, int PREC, int NW, int FPW, typename TIn, bool PIPE>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<LOG2M, NW>())))
void k_tile(TileArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr int M = 1 << LOG2M;
  constexpr int NI = (M / 2 + 1 + 63) / 64;       // split iterations per lane
  constexpr int TF = NW * FPW;                     // frames per tile
  constexpr int NT = NW * 64;
  constexpr bool SPEC = (MODE == OUT_STFT_MAG || MODE == OUT_STFT_POWER || MODE == OUT_STFT_COMPLEX);
  constexpr bool CPLX = (MODE == OUT_STFT_COMPLEX);
  constexpr int F = M + 1;
  constexpr int MAXB = kMaxMelBlocks / NW;         // mel block slots per wave (n_mels <= 256)
  static_assert(TF <= 16, "tile larger than the MFMA N dimension");

  const int lane0 = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int RB = a.lds_region_bytes;
  const int S = a.lds_row_stride;                  // staging row stride (elements)
  unsigned char* region = lds + wave * RB;
  float2* scr = reinterpret_cast<float2*>(region);
  constexpr int NB = PIPE ? 2 : 1;                           // staging / partial buffers
  int* sc_exp0 = reinterpret_cast<int*>(lds + NW * RB);      // [NB][16]
  float* red_max0 = reinterpret_cast<float*>(sc_exp0 + 16 * NB);   // [NB][NW]
  float* red_min0 = red_max0 + NW * NB;                       // [NB][NW]
  // the set of the tile being processed (PIPE: alternating per processed tile)
  int buf = 0;
  int* sc_exp = sc_exp0;
  float* red_max = red_max0;
  float* red_min = red_min0;
  unsigned char* stg = lds;                                   // its staging regions
  constexpr int ESZ = SPEC ? (CPLX ? 8 : 4) : (PREC == PREC_F32 ? 4 : 2);
  const int rowb = S * ESZ;
  // A wave's region: its FFT scratch with its last frame's staging row over
  // it (written once that frame's FFT is done), then the rows of its other
  // frames from max(scratch, row) on; FPW = 1 needs only the first part.
  // 16-bit mel / MFCC at n_fft 2048: the FFT runs on the matrix cores from
  // registers (mfft2048) and needs no LDS scratch
  constexpr bool MF = mfft_kernel(SPEC, PREC, LOG2M);
  constexpr int SCRB = MF ? 0 : ((M + M / 16) * (M >= 512 ? 4 : 8) + 15) / 16 * 16;
  const int row0 = max(SCRB, (rowb + 15) / 16 * 16);
  auto row_ptr = [&](int f) -> unsigned char* {   // staging row of tile frame f
    const int q = f % FPW;
    return stg + (f / FPW) * RB + ((q == FPW - 1) ? 0 : row0 + q * rowb);
  };

  const int t_begin = blockIdx.x * a.tiles_per_wg;
  const int t_end = min(t_begin + a.tiles_per_wg, a.n_tiles);
  if constexpr (MODE == OUT_MFCC) {
    // the next step's clip maxima (nobody reads them in this step)
    if (a.phase == 1 && a.clip_max_clear && blockIdx.x == 0)
      for (int c = threadIdx.x; c < a.n_clips; c += NT) a.clip_max_clear[c] = float_to_key(-INFINITY);
  }
  if constexpr (MODE == OUT_MFCC) {
    // phase 2 recomputes only tiles whose top_db floor is active: a
    // workgroup with none of them leaves before filling its LDS tables
    if (a.phase == 2) {
      bool act = false;
      if (a.use_top_db)
        for (int t = t_begin + (int)threadIdx.x; t < t_end; t += NW * 64)
          act |= a.tile_min[t] < key_to_float(a.clip_max_key[ldc(a.tile_clip, t)]) - a.top_db;
      // workgroup OR through one LDS flag per wave (its own slots, never
      // reused); the wave vote is taken with every lane active (a ballot
      // inside the lane-0 branch would see lane 0's tile only)
      int* any_act = reinterpret_cast<int*>(red_min0 + NW * NB);
      const bool wave_act = __builtin_amdgcn_ballot_w64(act) != 0;
      if ((threadIdx.x & 63) == 0) any_act[threadIdx.x >> 6] = wave_act;
      lds_barrier();
      int any = 0;
      for (int w = 0; w < NW; ++w) any |= any_act[w];
      if (!any) return;
    }
  }
  MM_PT_DECL
  // Samples of this wave's frame in the next tile, loaded while the current
  // tile's contraction and barriers run (register-split FFTs, one frame per
  // wave, interior frames with an even start; pf_tile = -1: none).
  using SHP = FftShape<LOG2M>;
  constexpr bool PF = SHP::REG_SPLIT && FPW == 1;
  float2 pf[SHP::Q0][SHP::R0];
  int pf_tile = -1;
#pragma unroll
  for (int q = 0; q < SHP::Q0; ++q)
#pragma unroll
    for (int r = 0; r < SHP::R0; ++r) pf[q][r] = make_float2(0.0f, 0.0f);
  // pf is written on every path, so it is dead between its use at the start
  // of a tile and the next prefetch (it never occupies registers in the FFT).
  // `r` is the record of tile_now + 1 (mask 0: nothing to fetch).
  auto prefetch_next = [&](int tile_now, const TilePf& r) {
    if constexpr (PF) {
      const bool ok = (r.mask >> wave) & 1u;
      pf_tile = ok ? tile_now + 1 : -1;
      if (ok) {
        const TIn* xf = wave_uniform(static_cast<const TIn*>(a.signal) + r.base + (int64_t)wave * a.hop);
        const int l = opaque(lane0);
        const unsigned lo = (unsigned)l * (2u * sizeof(TIn));
        if constexpr (MF) {   // the transform's pairs: mfft_point(lane, i)
          const unsigned lp = (unsigned)((l & 15) + 128 * (l >> 4)) * (2u * sizeof(TIn));
#pragma unroll
          for (int i = 0; i < 16; ++i) pf[0][i] = SamplePair<TIn>::load(xf + 2 * mfft_pair_off(i), lp);
        } else {
#pragma unroll
          for (int q = 0; q < SHP::Q0; ++q) {
            // lanes past the pass-0 butterflies (B0 < 64) hold no sample
            // pairs: their indices would run past the frame
            if ((SHP::B0 % 64 == 0) || l + 64 * q < SHP::B0) {
#pragma unroll
              for (int r2 = 0; r2 < SHP::R0; ++r2)   // sample pair n = lane + 64 q + r2 B0
                pf[q][r2] = SamplePair<TIn>::load(xf + 2 * (64 * q + r2 * SHP::B0), lo);
            } else {
#pragma unroll
              for (int r2 = 0; r2 < SHP::R0; ++r2) pf[q][r2] = make_float2(0.0f, 0.0f);
            }
          }
        }
      } else {
        // nothing to fetch: pf is unused on the next tile; an empty asm
        // definition ends the old values' lifetime without an instruction
#pragma unroll
        for (int q = 0; q < SHP::Q0; ++q)
#pragma unroll
          for (int r2 = 0; r2 < SHP::R0; ++r2) {
            __asm__ volatile("" : "=v"(pf[q][r2].x));
            __asm__ volatile("" : "=v"(pf[q][r2].y));
          }
      }
    }
  };
  if constexpr (PF) {
    if (t_begin < t_end && a.phase == 1) prefetch_next(t_begin - 1, ldc_pf(a.tile_pf, t_begin));
  }
  // LDS copies of the FFT tables (16-bit mel / MFCC, n_fft 1024 and 2048):
  // [window: n_fft floats][per-pass twiddles: M float2][W_N^k: M float2],
  // published by a barrier before the first FFT
  constexpr bool TABL = !MF && !SPEC && PREC != PREC_F32 && NW == 16 && SHP::REG_SPLIT && LOG2M <= 10;
  // the transform's table in LDS (any shape; mfft2048, with the window)
  const unsigned char* mft = nullptr;
  // (the pointers are set in a compile-time branch: a pointer that may be
  // either global or LDS would compile to flat loads)
  const float* win;
  const float2* twp;
  const float2* twn;
  if constexpr (MF) {
    unsigned char* tb = lds + a.tab_lds_off;
    if (t_begin < t_end) {
      for (int p = threadIdx.x; p < kMfTabBytes / 16; p += NT)
        reinterpret_cast<uint4*>(tb)[p] = reinterpret_cast<const uint4*>(a.mft)[p];
      // bins F .. S-1 of every staging row (both sets) read as zero in the
      // banded contraction; with no FFT scratch under the rows, staging
      // (bins 0 .. F-1) never touches them again
      const int npad = S - F;
      for (int idx = threadIdx.x; idx < NB * TF * npad; idx += NT) {
        const int rr = idx / npad, k = F + idx % npad;
        stg = lds + ((rr >= TF) ? a.lds_stage2_off : 0);
        reinterpret_cast<uint16_t*>(row_ptr(rr % TF))[k] = 0;
      }
      stg = lds;
      lds_barrier();
    }
    win = a.window;
    mft = tb;
    twp = a.twp;
    twn = a.twN;
  } else if constexpr (TABL) {
    unsigned char* tb = lds + a.tab_lds_off;
    if (t_begin < t_end) {
      const int nw4 = a.n_fft / 4, nt = M / 2;   // 16-byte pieces of the window / of each twiddle table
      for (int p = threadIdx.x; p < nw4 + 2 * nt; p += NT) {
        uint4 v;
        if (p < nw4) v = reinterpret_cast<const uint4*>(a.window)[p];
        else if (p < nw4 + nt) v = reinterpret_cast<const uint4*>(a.twp)[p - nw4];
        else v = reinterpret_cast<const uint4*>(a.twN)[p - nw4 - nt];
        reinterpret_cast<uint4*>(tb)[p] = v;
      }
      lds_barrier();
    }
    win = reinterpret_cast<const float*>(tb);
    twp = reinterpret_cast<const float2*>(tb + a.n_fft * 4);
    twn = twp + M;
  } else {
    win = a.window;
    twp = a.twp;
    twn = a.twN;
  }
  // LDS copies of the banded filterbank ([chunk][8 rows][80 B]: 64 B of weights per row) and of the
  // DCT operand tables; the first tile's barrier publishes them
  if constexpr (!SPEC) {
    if (a.fb_lds_off && t_begin < t_end) {
      using FE = std::conditional_t<PREC == PREC_F32, float, uint16_t>;
      const FE* tab;
      if constexpr (PREC == PREC_F32) tab = a.fb32;
      else tab = a.fb16;
      constexpr int EPP = 16 / (int)sizeof(FE);   // elements per 16-byte piece
      for (int p = threadIdx.x; p < a.cache_chunks * 32; p += NT) {
        const int c = p >> 5, r = (p >> 2) & 7, q = p & 3;
        const FE* src = tab + ldc(a.cache_src, c) + r * a.fbs + q * EPP;
        *reinterpret_cast<uint4*>(lds + a.fb_lds_off + c * kFbChunkBytes + r * kFbRowBytes + q * 16) =
            *reinterpret_cast<const uint4*>(src);
      }
    }
    if (MODE == OUT_MFCC && a.dct_lds_off && t_begin < t_end) {
      const int n16 = a.ndt * a.nmh * 64;   // 16-byte pieces per table (1 KB per (ct, block))
      if constexpr (PREC == PREC_F32) {
        for (int p = threadIdx.x; p < n16; p += NT)
          *reinterpret_cast<uint4*>(lds + a.dct_lds_off + p * 16) = reinterpret_cast<const uint4*>(a.dct32)[p];
      } else {
        const int nh = n16 / 2;   // hi and lo tables are 512 B per (ct, block)
        for (int p = threadIdx.x; p < n16; p += NT)
          *reinterpret_cast<uint4*>(lds + a.dct_lds_off + p * 16) =
              (p < nh) ? reinterpret_cast<const uint4*>(a.dct_hi)[p] : reinterpret_cast<const uint4*>(a.dct_lo)[p - nh];
      }
    }
  }
  // Top_db bookkeeping (thread 0) and the fixed-order block sum of a
  // finished MFCC tile whose DCT partials and per-wave extrema (set pb) a
  // barrier has published in LDS: thread t of the workgroup owns output
  // (coefficient 16 ct + (t >> 4) % 16, frame t & 15), so 16 consecutive
  // threads store 16 consecutive frames of one coefficient.
  auto finish_tile = [&](int ft, int pb) {
    const int fclip = ldc(a.tile_clip, ft), ff0 = ldc(a.tile_f0, ft), fT = ldc(a.clip_frames, fclip);
    const int po = PIPE ? (pb ? a.lds_part2_off : a.lds_part_off) : (a.lds_part_off > 0 ? a.lds_part_off : 0);
    const float* red = reinterpret_cast<const float*>(lds + po);
    if (a.use_top_db && a.phase == 1 && threadIdx.x == 0) {
      int bm = INT_MIN, bn = INT_MAX;
      for (int w = 0; w < NW; ++w) {
        bm = max(bm, reinterpret_cast<const int*>(red_max0 + NW * pb)[w]);
        bn = min(bn, reinterpret_cast<const int*>(red_min0 + NW * pb)[w]);
      }
      if (bm > INT_MIN) atomicMax(&a.clip_max_key[fclip], float_to_key(fkey_inv(bm)));
      a.tile_min[ft] = (bn < INT_MAX) ? fkey_inv(bn) : INFINITY;
    }
    const int nout = a.ndt * 256;
    for (int t = threadIdx.x; t < nout; t += NT) {
      const int ct = t >> 8, cl = (t >> 4) & 15, fi = t & 15;
      const int e = ((cl >> 2) * 16 + fi) * 4 + (cl & 3);   // MFMA layout: lane (cl / 4) * 16 + fi, item cl % 4
      const float* src = red + (int64_t)ct * a.nmh * 256 + e;
      float sm = 0.0f;
      int h = 0;
      for (; h + 8 <= a.nmh; h += 8) {   // 8 reads in flight, added in block order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[(h + u) * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) sm += v[u];
      }
      for (; h < a.nmh; ++h) sm += src[h * 256];
      const int c = 16 * ct + cl;
      if (fi < TF && ff0 + fi < fT && c < a.n_mfcc)
        reinterpret_cast<float*>(a.out)[ldc(a.out_off, fclip) + (int64_t)c * fT + ff0 + fi] = sm;
    }
  };
  int pend = -1, pbuf = 0;   // PIPE: the processed tile still to finish, and its set
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  for (int tile = t_begin; tile < t_end; ++tile) {
    // the prefetched samples are used first thing; on tiles that do not use
    // them nothing is in flight (see wait_vmem_all)
    if constexpr (PF) wait_vmem_all();
    // keep lane-derived table addresses inside the loop: hoisted out of it
    // they would stay live as 64-bit register pairs across the whole FFT
    int lane = lane0;
    __asm__ __volatile__("" : "+v"(lane));
    const int clip = ldc(a.tile_clip, tile);
    const int f0 = ldc(a.tile_f0, tile);
    const int T = ldc(a.clip_frames, clip);
    float thr = -INFINITY;
    if constexpr (MODE == OUT_MFCC) {
      if (a.use_top_db && a.phase == 2) {
        thr = key_to_float(a.clip_max_key[clip]) - a.top_db;
        if (!(a.tile_min[tile] < thr)) continue;   // clamp inactive: phase 1 result stands
      }
    }
    if constexpr (PIPE) {
      sc_exp = sc_exp0 + 16 * buf;
      red_max = red_max0 + NW * buf;
      red_min = red_min0 + NW * buf;
      stg = lds + (buf ? a.lds_stage2_off : 0);
    }

    // ---- FFT of this wave's frames; spectra kept in registers ----
    using SH = FftShape<LOG2M>;
    constexpr bool RS = SH::REG_SPLIT;
    constexpr int QL = RS ? SH::QL : 1, RL = RS ? SH::RL : 1;
    constexpr int NH = RS ? 1 : FPW;               // register-split frames hold nothing across FFTs
    constexpr int EC = CPLX ? 2 : 1;
    float hk[NH][NI][EC];
    float hm[NH][NI][EC];
    // register-split frames of a multi-frame wave are staged as soon as
    // their FFT is done (the next frame's FFT reuses the registers and the
    // scratch; earlier frames' rows lie outside the scratch)
    constexpr bool STAGE_EACH = RS && FPW > 1;
    // spectrum bin of register (b, r) after the paired real split: k for
    // r < RL/2, M - k of the pair (b, r - RL/2) above (lane 0, b = 0, r =
    // RL/2: M, the Nyquist bin)
    auto bin_of = [&](int b, int r) -> int {
      if constexpr (MF) {   // slot 2b + r of the transform (mfft_bin)
        if (r < RL / 2) return mfft_bin(lane, 2 * b + r);
        return M - mfft_bin(lane, 2 * b + r - RL / 2);
      }
      if (r < RL / 2) return lane + 64 * b + SH::BL * r;
      return M - (lane + 64 * b + SH::BL * (r - RL / 2));
    };
    constexpr int NQ = STAGE_EACH ? 1 : FPW;       // frames whose spectra are held at once
    float hv[NQ][QL][RL][EC];
    float hn[NQ][EC];                              // bin M/2 (valid in lane HL)
    constexpr int HL = MF ? 32 : 0;
    int fexp[NQ];
    int fadj[NQ];   // added to fexp for the contraction: undoes the transform's input scale
    // Per-frame power-of-two operand scale (16-bit contractions). A BF16
    // value has the FP32 exponent range, so scaling its operand by 2^-e and
    // the result by 2^e changes no bit unless a value leaves the normal
    // range; in the MFCC kernel anything that small lies below the amin floor
    // (1e-10) of the dB, so the BF16 MFCC kernel stages unscaled values. The
    // FP16 kernels (65504 at most) and the BF16 mel / chroma output keep it.
    constexpr bool FSCALE = !SPEC && (PREC == PREC_F16 || (PREC == PREC_BF16 && MODE != OUT_MFCC));
    // Writes frame q's spectrum (registers) into its staging row, and its
    // operand scale exponent for the contraction.
    auto put = [&](unsigned char* row, int k, const float* v, int e) {
      if constexpr (CPLX) {
        reinterpret_cast<float2*>(row)[k] = make_float2(v[0], v[EC - 1]);
      } else if constexpr (SPEC || PREC == PREC_F32) {
        reinterpret_cast<float*>(row)[k] = v[0];
      } else {
        const float sv = FSCALE ? ldexpf(v[0], -e) : v[0];
        reinterpret_cast<uint16_t*>(row)[k] = (PREC == PREC_BF16) ? f2bf(sv) : f2h(sv);
      }
    };
    auto stage = [&](int q) {
      const int qh = STAGE_EACH ? 0 : q;
      const int fl = wv * FPW + q;
      const int fr = f0 + fl;
      unsigned char* row = row_ptr(fl);
      if (fr < T) {
        if constexpr (MF) {
          // the transform's bins (mfft_bin): slot (p, j) of this lane at
          // byte ob[p] + 128 j of the row and its mirror at 2048 - ob[p] -
          // 128 j; p = 0 and p = 1 of one slot share a packed conversion
          const int lb = 4 * (lane & 15) + 512 * (lane >> 4);
          const int ob0 = lb + (mfft_own_odd(lane, 0) ? 64 : 0), ob1 = lb + 2;
          auto cvt2 = [&](float x0, float x1) -> unsigned {
            if constexpr (FSCALE) {
              x0 = ldexpf(x0, -fexp[qh]);
              x1 = ldexpf(x1, -fexp[qh]);
            }
            if constexpr (PREC == PREC_BF16) {
              typedef __bf16 b2v_ __attribute__((ext_vector_type(2)));
              return __builtin_bit_cast(unsigned, __builtin_convertvector((f2v){x0, x1}, b2v_));
            } else {
              typedef _Float16 h2v_ __attribute__((ext_vector_type(2)));
              return __builtin_bit_cast(unsigned, __builtin_convertvector((f2v){x0, x1}, h2v_));
            }
          };
#pragma unroll
          for (int j = 0; j < 4; ++j) {   // register (b, r) of slot 4p + j: b = 2p + j / 2, r = j % 2 (+2: mirror)
            const unsigned o = cvt2(hv[qh][j >> 1][j & 1][0], hv[qh][2 + (j >> 1)][j & 1][0]);
            const unsigned m = cvt2(hv[qh][j >> 1][2 + (j & 1)][0], hv[qh][2 + (j >> 1)][2 + (j & 1)][0]);
            *reinterpret_cast<uint16_t*>(row + ob0 + 128 * j) = (uint16_t)o;
            *reinterpret_cast<uint16_t*>(row + ob1 + 128 * j) = (uint16_t)(o >> 16);
            *reinterpret_cast<uint16_t*>(row + 2048 - ob0 - 128 * j) = (uint16_t)m;
            *reinterpret_cast<uint16_t*>(row + 2048 - ob1 - 128 * j) = (uint16_t)(m >> 16);
          }
          if (lane == HL) put(row, M / 2, hn[qh], fexp[qh]);
        } else if constexpr (RS) {
#pragma unroll
          for (int b = 0; b < QL; ++b)
#pragma unroll
            for (int r = 0; r < RL; ++r) put(row, bin_of(b, r), hv[qh][b][r], fexp[qh]);
          if (lane == HL) put(row, M / 2, hn[qh], fexp[qh]);
        } else {
#pragma unroll
          for (int i = 0; i < NI; ++i) {
            const int k = lane + 64 * i;
            if ((M / 2 + 1) % 64 == 0 || k <= M / 2) {
              put(row, k, hk[qh][i], fexp[qh]);
              put(row, M - k, hm[qh][i], fexp[qh]);
            }
          }
        }
        if constexpr (!SPEC && !MF) {
          // bins F .. S-1 must read as zero in the banded contraction (the
          // FFT scratch under the row overwrites them every tile; the
          // matrix-core kernels zero them once)
          for (int k = F + lane; k < S; k += 64) {
            if constexpr (PREC == PREC_F32) reinterpret_cast<float*>(row)[k] = 0.0f;
            else reinterpret_cast<uint16_t*>(row)[k] = 0;
          }
        }
      }
      if constexpr (!SPEC) {
        if (lane == 0) sc_exp[fl] = (fr < T) ? fexp[qh] + fadj[qh] : 0;
      }
    };
    auto frame_body = [&](int q) {
      const int qh = STAGE_EACH ? 0 : q;
      fexp[qh] = 0;
      fadj[qh] = 0;
      const int fr = f0 + wv * FPW + q;   // wave-uniform: framing math stays scalar
      if (fr < T) {
        float vmax = 0.0f;
        if constexpr (RS) {
          // hv[b][r]: 2 X[k] of pair (b, r) for r < RL/2, 2 X[M-k] of pair
          // (b, r - RL/2) for r >= RL/2 (see bin_of); hn: X[M/2] in lane HL
          float2 xk[QL][RL / 2], xm[QL][RL / 2];
          float2 xh;
          int es = 0;   // the transform's input scale (2^es)
          if constexpr (MF) {
            float2 v[16], k8[8], m8[8];
            mfft_load<TIn>(a, clip, fr, lane, mft, v, pf[0], PF && pf_tile == tile);
            es = mfft_frame(v, mft, lane, k8, m8, xh);
#pragma unroll
            for (int b = 0; b < QL; ++b)
#pragma unroll
              for (int r = 0; r < RL / 2; ++r) {
                xk[b][r] = k8[2 * b + r];
                xm[b][r] = m8[2 * b + r];
              }
          } else {
            float2 z[QL][RL];
            frame_fft_split<LOG2M, TIn>(a, clip, fr, reinterpret_cast<float*>(scr), lane, win, twp, z, pf,
                                        PF && pf_tile == tile);
            real_split_pairs<LOG2M>(z, twn, lane, xk, xm, xh);
          }
          auto set = [&](float* h, const float2 X) {   // X = 2 X[bin]
            if constexpr (CPLX) {
              h[0] = 0.5f * X.x;
              h[EC - 1] = 0.5f * X.y;
            } else {
              const float p2 = fmaf(X.x, X.x, X.y * X.y);   // 4 |X|^2
              float v = p2;
              if constexpr (MODE == OUT_STFT_MAG) v = 0.5f * sqrtf(p2);
              else if constexpr (MODE == OUT_STFT_POWER) v = 0.25f * p2;
              h[0] = v;
            }
          };
#pragma unroll
          for (int b = 0; b < QL; ++b)
#pragma unroll
            for (int r = 0; r < RL / 2; ++r) {
              set(hv[qh][b][r], xk[b][r]);
              set(hv[qh][b][r + RL / 2], xm[b][r]);
            }
          set(hn[qh], xh);
          if constexpr (!SPEC) {
            if constexpr (MF) {
              // the input scale 2^es: folded into the frame exponent for
              // powers 2 and 1, removed before the power otherwise
              if (es != 0) {
                if (a.power == 2.0f) {
                  fadj[qh] = -2 * es;
                } else if (a.power == 1.0f) {
                  fadj[qh] = -es;
                } else {
#pragma unroll
                  for (int b = 0; b < QL; ++b)
#pragma unroll
                    for (int r = 0; r < RL; ++r) hv[qh][b][r][0] = ldexpf(hv[qh][b][r][0], -2 * es);
                  hn[qh][0] = ldexpf(hn[qh][0], -2 * es);
                }
              }
            }
            // (2|X|)^p from (2|X|)^2 (the 2^p is removed after the mel
            // product): one wave-uniform branch per frame, so the default
            // power 2 runs no transcendental instruction at all
            if (a.power != 2.0f) {
              if (a.power == 1.0f) {
#pragma unroll
                for (int b = 0; b < QL; ++b)
#pragma unroll
                  for (int r = 0; r < RL; ++r) hv[qh][b][r][0] = sqrtf(hv[qh][b][r][0]);
                hn[qh][0] = sqrtf(hn[qh][0]);
              } else {
#pragma unroll
                for (int b = 0; b < QL; ++b)
#pragma unroll
                  for (int r = 0; r < RL; ++r) hv[qh][b][r][0] = pow_from_p2(hv[qh][b][r][0], a.power);
                hn[qh][0] = pow_from_p2(hn[qh][0], a.power);
              }
            }
#pragma unroll
            for (int b = 0; b < QL; ++b)
#pragma unroll
              for (int r = 0; r < RL; ++r) vmax = fmaxf(vmax, hv[qh][b][r][0]);
            if (lane == HL) vmax = fmaxf(vmax, hn[qh][0]);
          }
        } else {
          frame_fft<LOG2M, TIn>(a, clip, fr, scr, lane);
#pragma unroll
          for (int i = 0; i < NI; ++i) {
            const int k = lane + 64 * i;
            if ((M / 2 + 1) % 64 == 0 || k <= M / 2) {
              float2 xk, xmk;
              real_split<LOG2M>(scr, a.twN, k, xk, xmk);
              if constexpr (CPLX) {
                hk[qh][i][0] = xk.x; hk[qh][i][EC - 1] = xk.y;
                hm[qh][i][0] = xmk.x; hm[qh][i][EC - 1] = xmk.y;
              } else {
                const float pk = fmaf(xk.x, xk.x, xk.y * xk.y);
                const float pm = fmaf(xmk.x, xmk.x, xmk.y * xmk.y);
                if constexpr (MODE == OUT_STFT_MAG) {
                  hk[qh][i][0] = sqrtf(pk);
                  hm[qh][i][0] = sqrtf(pm);
                } else if constexpr (MODE == OUT_STFT_POWER) {
                  hk[qh][i][0] = pk;
                  hm[qh][i][0] = pm;
                } else {
                  // same 2^p scale as the register split (removed after the mel product)
                  hk[qh][i][0] = pow_from_p2(4.0f * pk, a.power);
                  hm[qh][i][0] = pow_from_p2(4.0f * pm, a.power);
                  vmax = fmaxf(vmax, fmaxf(hk[qh][i][0], hm[qh][i][0]));
                }
              }
            }
          }
        }
        wave_lds_fence();
        if constexpr (FSCALE) {
          // per-frame power-of-two scale so the 16-bit operand keeps the
          // frame's top range (exactly undone after the contraction): bf16
          // keeps the FP32 exponent range, FP16 maps the frame maximum into
          // [2^14, 2^15) so ~84 dB below it stay normal numbers
          vmax = wave_max_nonneg(vmax);
          int e = 0;
          if (vmax > 0.0f && vmax <= 3.0e38f) e = __builtin_amdgcn_frexp_expf(vmax) - 1 - (PREC == PREC_F16 ? 14 : 0);
          fexp[qh] = e;
        }
      }
      if constexpr (STAGE_EACH) {
        wave_lds_fence();
        stage(q);
        wave_lds_fence();
      }
    };
    if constexpr (STAGE_EACH) {
      // one frame at a time (a rolled loop: nothing of one frame's FFT is
      // scheduled into the other's)
#pragma unroll 1
      for (int q = 0; q < FPW; ++q) frame_body(q);
    } else {
#pragma unroll
      for (int q = 0; q < FPW; ++q) frame_body(q);
    }

    MM_PT(0);
    // Scalar loads for after the barrier, issued now so they have landed by
    // then (loaded at the top of the tile they would occupy scalar registers
    // through the FFT): the next tile's prefetch record and this wave's mel
    // block slots (block | first bin << 6 | chunks << 19; -1: empty).
    TilePf nrec;
    nrec.base = 0;
    nrec.mask = 0u;
    if constexpr (PF) {
      if (tile + 1 < t_end && a.phase == 1) nrec = ldc_pf(a.tile_pf, tile + 1);
    }
    int slot[MAXB];
    if constexpr (!SPEC) {
      const int wq = opaque_s(wv);   // keeps the loop-invariant loads inside the loop
#pragma unroll
      for (int bi = 0; bi < MAXB; ++bi) slot[bi] = ldc(a.assign, wq * MAXB + bi);
    }
    // ---- staging rows (this wave's region; its scratch is no longer used) ----
    if constexpr (!STAGE_EACH) {
#pragma unroll
      for (int q = 0; q < FPW; ++q) stage(q);
    }
    // Matrix-core FFT kernels: the next tile's samples are requested as soon
    // as this wave's frame is staged. Waves on one SIMD finish their FFTs
    // thousands of cycles apart, so the workgroup's 16 requests reach the
    // vector memory path spread out instead of in one burst after the
    // contraction (where each wave stalled issuing them); the samples stay in
    // registers through the contraction, which needs few.
    if constexpr (MF) prefetch_next(tile, nrec);
    MM_PT(1);
    lds_barrier();
    MM_PT(2);
    if constexpr (PIPE) {
      if (pend >= 0) finish_tile(pend, pbuf);   // its partials were published by this barrier
    }

    if constexpr (SPEC) {
      // ---- transposed write-out: [F, T], runs of TF frames per bin ----
      const int nf = min(TF, T - f0);
      const int64_t base = ldc(a.out_off, clip);
      for (int idx = threadIdx.x; idx < F * TF; idx += NT) {
        const int f = idx % TF;
        const int k = idx / TF;
        if (f < nf) {
          const unsigned char* row = row_ptr(f);
          if constexpr (CPLX) {
            reinterpret_cast<float2*>(a.out)[base + (int64_t)k * T + f0 + f] =
                reinterpret_cast<const float2*>(row)[k];
          } else {
            reinterpret_cast<float*>(a.out)[base + (int64_t)k * T + f0 + f] =
                reinterpret_cast<const float*>(row)[k];
          }
        }
      }
      prefetch_next(tile, nrec);
      lds_barrier();
      continue;
    } else {
      // ---- banded mel contraction on MFMA ----
      // Blocks of 8 mels: MFMA rows 8..15 repeat rows 0..7 (each output row
      // depends on its own A row only, and rows 8..15 are discarded), which
      // halves the useful MFMA rate but lets the slowest wave of a tile hold
      // ~40% fewer bin chunks than blocks of 16 would (MFMA time is not the
      // bound). Filterbank and DCT operands come from their LDS copies when
      // the launch has them (fb_lds_off / dct_lds_off), else from L2.
      const int ri = lane & 15;          // MFMA row/col index within 16
      const int g = lane >> 4;           // lane group (K slice)
      // columns past an 8-frame tile read row 0 and are never stored
      const unsigned char* prow = row_ptr(ri < TF ? ri : 0);
      const int myexp = (ri < TF) ? sc_exp[ri] : 0;
      const bool frame_ok = (ri < TF) && (f0 + ri) < T;
      const float vscale = (PREC == PREC_F32) ? a.post_scale : ldexpf(a.post_scale, myexp);
      const bool track = (MODE == OUT_MFCC) && a.use_top_db && a.phase == 1;   // phase 1: tile extrema for the clamp
      f32x4 dbv[MAXB];
      int kmax = INT_MIN, kmin = INT_MAX;   // fkey of this lane's dB extrema
      int nb = 0;   // slots in use (filled as a prefix)
#pragma unroll
      for (int bi = 0; bi < MAXB; ++bi)
        if (slot_used(slot[bi])) nb = bi + 1;
      // DCT-II operands of (row tile ct, mel block h): A entries for this lane
      struct DctOp { float f[4]; uint2 hi, lo; };
      auto load_dct = [&](int ct, int h) -> DctOp {
        DctOp o;
        const int e0 = (ct * a.nmh + h) * 256;   // entries of (ct, h)
        if (a.dct_lds_off) {   // f32 [..][4][64]; 16-bit hi block then lo block, [..][64][4]
          if constexpr (PREC == PREC_F32) {
            const float* d = reinterpret_cast<const float*>(lds + a.dct_lds_off) + e0;
#pragma unroll
            for (int j = 0; j < 4; ++j) o.f[j] = d[j * 64 + lane];
          } else {
            const uint16_t* d = reinterpret_cast<const uint16_t*>(lds + a.dct_lds_off);
            o.hi = *reinterpret_cast<const uint2*>(d + e0 + lane * 4);
            o.lo = *reinterpret_cast<const uint2*>(d + a.ndt * a.nmh * 256 + e0 + lane * 4);
          }
        } else if constexpr (PREC == PREC_F32) {
#pragma unroll
          for (int j = 0; j < 4; ++j) o.f[j] = ldg<float>(a.dct32 + e0 + j * 64, (unsigned)lane * 4u);
        } else {
          o.hi = ldg<uint2>(a.dct_hi + e0, (unsigned)lane * 8u);
          o.lo = ldg<uint2>(a.dct_lo + e0, (unsigned)lane * 8u);
        }
        return o;
      };
      using AV = std::conditional_t<PREC == PREC_F32, float4, uint4>;   // one chunk, one lane
      using FE = std::conditional_t<PREC == PREC_F32, float, uint16_t>;
      constexpr int CH = (PREC == PREC_F32) ? 16 : 32;                  // bins per chunk
      auto mfma_chunk = [&](const AV& av, const AV& bv, f32x4 acc) -> f32x4 {
        if constexpr (PREC == PREC_F32) {
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc, 0, 0, 0);
        } else if constexpr (PREC == PREC_BF16) {
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                        acc, 0, 0, 0);
        } else {
          acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, av), __builtin_bit_cast(f16x8, bv),
                                                       acc, 0, 0, 0);
        }
        return acc;
      };
      // mel^T of the block in slot v over its bin band, in rounds of 8, 4, 2
      // or 1 chunks (chosen by the chunks left): a round issues all its
      // operand loads before its MFMAs, so their latencies overlap, and loads
      // at most 3 chunks twice (clamped to the last chunk, MFMA skipped)
      auto mel_block = [&](int v) -> f32x4 {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const int h = slot_block(v), x0 = slot_first_bin(v), n = slot_chunks(v);
        const FE* pr = reinterpret_cast<const FE*>(prow) + x0 + (CH / 4) * g;
        // this lane's filterbank operand: row ri & 7, bins (CH/4) g .. of a chunk
        const bool cached = a.fb_lds_off != 0 && slot_cache_first(v) + n <= a.cache_chunks;
        const unsigned char* fbl = lds + a.fb_lds_off + slot_cache_first(v) * kFbChunkBytes + (ri & 7) * kFbRowBytes + g * 16;
        const FE* fbtab;
        if constexpr (PREC == PREC_F32) fbtab = a.fb32;
        else fbtab = a.fb16;
        const FE* fbb = fbtab + (int64_t)kMelBlock * h * a.fbs + x0;
        const unsigned fo = (unsigned)(((ri & (kMelBlock - 1)) * a.fbs + (CH / 4) * g) * (int)sizeof(FE));
        auto round = [&](auto kc, int c) {
          constexpr int K = decltype(kc)::value;
          AV av[K], bv[K];
          if (cached) {   // LDS copy: [chunk][8 rows][80 B]
#pragma unroll
            for (int u = 0; u < K; ++u) av[u] = *reinterpret_cast<const AV*>(fbl + kFbChunkBytes * min(c + u, n - 1));
          } else {
#pragma unroll
            for (int u = 0; u < K; ++u) av[u] = ldg<AV>(fbb + CH * min(c + u, n - 1), fo);
          }
#pragma unroll
          for (int u = 0; u < K; ++u) bv[u] = *reinterpret_cast<const AV*>(pr + CH * min(c + u, n - 1));
#pragma unroll
          for (int u = 0; u < K; ++u)
            if (u == 0 || c + u < n) acc = mfma_chunk(av[u], bv[u], acc);
          mfma_settle(acc);
          // duplicate loads whose MFMA was skipped end here (operands from
          // L2 only: an unconditional wait would also hold the LDS-operand
          // rounds until the next tile's sample prefetch has landed)
          if (!cached) wait_vmem_all();
        };
        for (int c = 0; c < n;) {
          const int rem = n - c;
          // rounds of at most 4 chunks in the matrix-core FFT kernels: they
          // hold the next tile's samples in registers through the contraction
          if (!MF && rem > 4) { round(std::integral_constant<int, 8>{}, c); c += 8; }
          else if (rem > 2) { round(std::integral_constant<int, 4>{}, c); c += 4; }
          else if (rem == 2) { round(std::integral_constant<int, 2>{}, c); c += 2; }
          else { round(std::integral_constant<int, 1>{}, c); c += 1; }
        }
        return acc;
      };
      // acc (of block h) -> mel values, then the MEL store or dB
      auto finish_block = [&](int h, int bi, f32x4 acc) {
        // acc[j] = mel[8h + 4(g & 1) + j] of tile frame ri (lane groups 2, 3
        // repeat groups 0, 1); vscale undoes the frame's power-of-two operand
        // scale and the 2^p of the contraction in one exact multiply
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] *= vscale;
        if constexpr (MODE == OUT_MEL) {
          if (frame_ok && g < 2) {
            float* o = reinterpret_cast<float*>(a.out) + ldc(a.out_off, clip);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int m = kMelBlock * h + 4 * g + j;
              if (m < a.n_mels) o[(int64_t)m * T + f0 + ri] = acc[j];
            }
          }
        } else {
          // 10 log10(max(mel, amin) / ref) = 10 log10(2) log2(max(mel, amin))
          // - 10 log10(ref) (db_off = 0 for the default ref = 1, where this
          // is the same float as the quotient form). The DCT table is zero in
          // the K slots of lane groups 2, 3 and of mels past n_mels, so those
          // slots need no masking.
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float d = fmaf(3.0102999566398120f, __builtin_amdgcn_logf(fmaxf(acc[j], a.amin)), -a.db_off);
            dbv[bi][j] = fmaxf(d, thr);   // thr = -inf except in phase 2 (floor at the clip max - top_db)
          }
          if (track) {
            if (kMelBlock * h + kMelBlock <= a.n_mels) {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                kmax = max(kmax, fkey(dbv[bi][j]));
                kmin = min(kmin, fkey(dbv[bi][j]));
              }
            } else {   // the last block of an n_mels that is not a multiple of 8
#pragma unroll
              for (int j = 0; j < 4; ++j)
                if (kMelBlock * h + 4 * (g & 1) + j < a.n_mels) {
                  kmax = max(kmax, fkey(dbv[bi][j]));
                  kmin = min(kmin, fkey(dbv[bi][j]));
                }
            }
          }
        }
      };
#pragma unroll
      for (int bi = 0; bi < MAXB; ++bi) {
        if (bi < nb) {
          const int v = slot[bi];
          finish_block(slot_block(v), bi, mel_block(v));
        }
      }
      // waves without mel blocks reach this point at once: their prefetch
      // overlaps the other waves' contraction
      if constexpr (MODE == OUT_MEL && !MF) prefetch_next(tile, nrec);
      MM_PT(3);

      if constexpr (MODE == OUT_MFCC) {
        if (track) {
          // lanes of frames past the clip end hold stale rows: neutral values
          const int tmax = wave_reduce_i<true>(frame_ok ? kmax : INT_MIN);
          const int tmin = wave_reduce_i<false>(frame_ok ? kmin : INT_MAX);
          if (lane == 0) {
            reinterpret_cast<int*>(red_max)[wave] = tmax;
            reinterpret_cast<int*>(red_min)[wave] = tmin;
          }
        }
        MM_PT(4);
        // DCT partials, one per (row tile ct, mel block h): [ndt][nmh][64
        // lanes][4], summed over h in increasing order below, so the result
        // does not depend on which wave contracted which block (any
        // workgroup shape gives the same bits). In their own LDS area when
        // the layout has one (2 barriers per tile), else over the staging
        // regions once every wave has read them (4 barriers).
        const bool sep = a.lds_part_off > 0;
        float* red = reinterpret_cast<float*>(
            lds + (PIPE ? (buf ? a.lds_part2_off : a.lds_part_off) : (sep ? a.lds_part_off : 0)));
        if (!sep) lds_barrier();   // staging reads done; region memory is reused below
        for (int ct = 0; ct < a.ndt; ++ct) {
#pragma unroll
          for (int bi = 0; bi < MAXB; ++bi) {
            if (bi < nb) {
              const int h = slot_block(slot[bi]);
              const DctOp op = load_dct(ct, h);
              f32x4 accd = {0.0f, 0.0f, 0.0f, 0.0f};
              if constexpr (PREC == PREC_F32) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  accd = __builtin_amdgcn_mfma_f32_16x16x4f32(op.f[j], dbv[bi][j], accd, 0, 0, 0);
              } else {
                // hi / lo split of the dB values (two conversions per pair)
                uint2 bhi2, blo2;
                split16<PREC>(dbv[bi][0], dbv[bi][1], bhi2.x, blo2.x);
                split16<PREC>(dbv[bi][2], dbv[bi][3], bhi2.y, blo2.y);
                if constexpr (PREC == PREC_BF16) {
                  const s16x4 Ah = __builtin_bit_cast(s16x4, op.hi), Al = __builtin_bit_cast(s16x4, op.lo);
                  const s16x4 Bh = __builtin_bit_cast(s16x4, bhi2), Bl = __builtin_bit_cast(s16x4, blo2);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bh, accd, 0, 0, 0);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al, Bh, accd, 0, 0, 0);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bl, accd, 0, 0, 0);
                } else {
                  const f16x4 Ah = __builtin_bit_cast(f16x4, op.hi), Al = __builtin_bit_cast(f16x4, op.lo);
                  const f16x4 Bh = __builtin_bit_cast(f16x4, bhi2), Bl = __builtin_bit_cast(f16x4, blo2);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16f16(Ah, Bh, accd, 0, 0, 0);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16f16(Al, Bh, accd, 0, 0, 0);
                  accd = __builtin_amdgcn_mfma_f32_16x16x16f16(Ah, Bl, accd, 0, 0, 0);
                }
              }
              mfma_settle(accd);
              float* dst = red + ((int64_t)(ct * a.nmh + h) * 64 + lane) * 4;
              *reinterpret_cast<float4*>(dst) = make_float4(accd[0], accd[1], accd[2], accd[3]);
            }
          }
        }
        MM_PT(5);
        if constexpr (!MF) prefetch_next(tile, nrec);
        MM_PT(6);
        if constexpr (PIPE) {
          pend = tile;   // finished after the next tile's barrier
          pbuf = buf;
          buf ^= 1;
          continue;
        }
        lds_barrier();   // partials and per-wave extrema are in LDS
        MM_PT(7);
        finish_tile(tile, 0);
        MM_PT(8);
        if (sep) continue;
      }
      lds_barrier();
    }
  }
  if constexpr (PIPE) {
    if (pend >= 0) {
      lds_barrier();
      finish_tile(pend, pbuf);
    }
  }
  MM_PT_STORE();
}

// ---------------------------------------------------------------------------
// generic path: any power-of-two n_fft up to 16384, one clip, plain kernels
// ---------------------------------------------------------------------------

// One workgroup per frame: radix-2 Stockham on M complex points in LDS
// (ping-pong), then the real split. Writes frame-major spectra spec[t][F]
// (complex) into a scratch buffer.
__global__ __launch_bounds__(256) void k_generic_fft(const float* __restrict__ x, int64_t len, int n_fft,
                                                     int hop, int center, int pad_mode,
                                                     const float* __restrict__ window,
                                                     const float2* __restrict__ tw,
                                                     const float2* __restrict__ twN, int n_frames,
                                                     float2* __restrict__ spec) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int fr = blockIdx.x;
  if (fr >= n_frames) return;
  const int F = n_fft / 2 + 1;
  const int64_t s0 = (int64_t)fr * hop - (center ? (n_fft / 2) : 0);
  float2* out = spec + (int64_t)fr * F;
  if (n_fft == 1) {
    if (threadIdx.x == 0) {
      const int64_t i0 = pad_index(s0, len, pad_mode);
      out[0] = make_float2((i0 >= 0 ? x[i0] : 0.0f) * window[0], 0.0f);
    }
    return;
  }
  const int M = n_fft / 2;
  float2* A = reinterpret_cast<float2*>(lds);
  float2* B = A + M;
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    const int64_t i0 = pad_index(s0 + 2 * n, len, pad_mode);
    const int64_t i1 = pad_index(s0 + 2 * n + 1, len, pad_mode);
    const float x0 = (i0 >= 0) ? x[i0] : 0.0f, x1 = (i1 >= 0) ? x[i1] : 0.0f;
    A[n] = make_float2(x0 * window[2 * n], x1 * window[2 * n + 1]);
  }
  __syncthreads();
  // radix-2 Stockham: Ns = 1, 2, 4, ...
  for (int Ns = 1; Ns < M; Ns <<= 1) {
    for (int j = threadIdx.x; j < M / 2; j += blockDim.x) {
      const int k = j % Ns;
      float2 u = A[j], v = A[j + M / 2];
      v = cmul(v, tw[k * (M / (2 * Ns))]);   // W_{2Ns}^k
      const int base = (j / Ns) * Ns * 2 + k;
      B[base] = cadd(u, v);
      B[base + Ns] = csub(u, v);
    }
    __syncthreads();
    float2* t = A; A = B; B = t;
  }
  for (int k = threadIdx.x; k <= M / 2; k += blockDim.x) {
    const float2 Az = A[k];
    const float2 Bz = A[(M - k) % M];
    const float2 Bc = make_float2(Bz.x, -Bz.y);
    const float2 E = make_float2(0.5f * (Az.x + Bc.x), 0.5f * (Az.y + Bc.y));
    const float2 O = make_float2(0.5f * (Az.x - Bc.x), 0.5f * (Az.y - Bc.y));
    const float2 T = cmul(O, twN[k]);
    out[k] = make_float2(E.x + T.y, E.y - T.x);
    out[M - k] = make_float2(E.x - T.y, -(E.y + T.x));
  }
}

// spec [T][F] complex -> out [F, T] in the requested STFT form
__global__ void k_generic_spec_out(const float2* __restrict__ spec, int F, int T, int kind, void* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)F * T) return;
  const int k = (int)(i / T), t = (int)(i % T);
  const float2 X = spec[(int64_t)t * F + k];
  if (kind == OUT_STFT_COMPLEX) reinterpret_cast<float2*>(out)[i] = X;
  else if (kind == OUT_STFT_POWER) reinterpret_cast<float*>(out)[i] = fmaf(X.x, X.x, X.y * X.y);
  else reinterpret_cast<float*>(out)[i] = sqrtf(fmaf(X.x, X.x, X.y * X.y));
}

// mel[m][t] = sum_k FB[m][k] * |X[t][k]|^power (dense, plain FP32)
__global__ void k_generic_mel(const float2* __restrict__ spec, int F, int T, const float* __restrict__ fb,
                              int fbs, int n_mels, float power, float* __restrict__ mel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_mels * T) return;
  const int m = (int)(i / T), t = (int)(i % T);
  const float2* X = spec + (int64_t)t * F;
  const float* w = fb + (int64_t)m * fbs;
  float s = 0.0f;
  for (int k = 0; k < F; ++k) {
    const float wk = w[k];
    if (wk != 0.0f) s = fmaf(wk, pow_from_p2(fmaf(X[k].x, X[k].x, X[k].y * X[k].y), power), s);
  }
  mel[i] = s;
}

// dB conversion with a running maximum (ordered-int atomic)
__global__ void k_db(const float* __restrict__ in, int64_t n, float mult, float ref, float amin,
                     float* __restrict__ out, unsigned* max_key) {
  float lmax = -INFINITY;
  auto db = [&](float x) {
    const float d = mult * log10f(fmaxf(x, amin) / ref);
    lmax = fmaxf(lmax, d);
    return d;
  };
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0) {
    // 16-byte loads and stores, then the tail
    const int64_t n4 = n / 4;
    for (int64_t i = t0; i < n4; i += stride) {
      const float4 v = reinterpret_cast<const float4*>(in)[i];
      reinterpret_cast<float4*>(out)[i] = make_float4(db(v.x), db(v.y), db(v.z), db(v.w));
    }
    for (int64_t i = 4 * n4 + t0; i < n; i += stride) out[i] = db(in[i]);
  } else {
    for (int64_t i = t0; i < n; i += stride) out[i] = db(in[i]);
  }
  // one atomic per block (same-address atomics serialise in L2): wave
  // maxima, then the block's 4 waves through LDS
  __shared__ float red[4];
  lmax = wave_max(lmax);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (m > -INFINITY) atomicMax(max_key, float_to_key(m));
  }
}

__global__ void k_db_clamp(float* __restrict__ x, int64_t n, const unsigned* max_key, float top_db) {
  const float thr = key_to_float(*max_key) - top_db;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = fmaxf(x[i], thr);
}

// out[c][t] = sum_m D[c][m] * logmel[m][t]
__global__ void k_generic_dct(const float* __restrict__ logmel, int n_mels, int T, const float* __restrict__ dct,
                              int n_mfcc, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_mfcc * T) return;
  const int c = (int)(i / T), t = (int)(i % T);
  float s = 0.0f;
  for (int m = 0; m < n_mels; ++m) s = fmaf(dct[(int64_t)c * n_mels + m], logmel[(int64_t)m * T + t], s);
  out[i] = s;
}

__global__ void k_fill_u32(unsigned* p, int64_t n, unsigned v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// table generation in FP64 on the device, rounded once
// ---------------------------------------------------------------------------

__device__ inline double hz_to_mel_d(double hz) {
  const double f_sp = 200.0 / 3.0, min_log_hz = 1000.0, min_log_mel = min_log_hz / f_sp;
  const double logstep = log(6.4) / 27.0;
  return (hz >= min_log_hz) ? min_log_mel + log(hz / min_log_hz) / logstep : hz / f_sp;
}
__device__ inline double mel_to_hz_d(double mel) {
  const double f_sp = 200.0 / 3.0, min_log_hz = 1000.0, min_log_mel = min_log_hz / f_sp;
  const double logstep = log(6.4) / 27.0;
  return (mel < min_log_mel) ? mel * f_sp : min_log_hz * exp((mel - min_log_mel) * logstep);
}

__global__ void k_tab_window(float* w, int n_fft, int win) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_fft) return;
  int m;
  if (win < n_fft) m = n - (n_fft - win) / 2;
  else m = n + (win - n_fft) / 2;
  double v = 0.0;
  if (win == 1) v = (m == 0) ? 1.0 : 0.0;
  else if (win > 1 && m >= 0 && m < win) v = 0.5 - 0.5 * cospi(2.0 * (double)m / (double)win);
  w[n] = (float)v;
}

__global__ void k_tab_twiddle(float2* tw, int M, float2* twN, int n_fft) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < M) tw[t] = make_float2((float)cospi(2.0 * t / M), (float)(-sinpi(2.0 * t / M)));
  if (t < M || t == 0) twN[t] = make_float2((float)cospi(2.0 * t / n_fft), (float)(-sinpi(2.0 * t / n_fft)));
}

// the matrix-core FFT table (mfft_table_piece), one 16-byte piece per thread
__global__ void k_tab_mfft(const float* win, unsigned char* tab) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMfTabBytes / 16) mfft_table_store(p, win, tab);
}

// entry [(r-1) NS + k] = W_{NS R}^{k r} of one Stockham pass
__global__ void k_tab_twiddle_pass(float2* twp, int NS, int R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (R - 1) * NS) return;
  const int r = i / NS + 1, k = i % NS;
  const double e = 2.0 * (double)k * (double)r / ((double)NS * (double)R);
  twp[i] = make_float2((float)cospi(e), (float)(-sinpi(e)));
}

__global__ void k_tab_mel(float* fb32, uint16_t* fbb, uint16_t* fbh, int fbs, int rows, int sr, int n_fft,
                          int n_mels, double fmin, double fmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * fbs) return;
  const int m = (int)(i / fbs), k = (int)(i % fbs);
  const int F = n_fft / 2 + 1;
  double w = 0.0;
  if (m < n_mels && k < F) {
    const double mlo = hz_to_mel_d(fmin), mhi = hz_to_mel_d(fmax);
    const double step = (mhi - mlo) / (double)(n_mels + 1);
    const double fl = mel_to_hz_d(mlo + m * step), fc = mel_to_hz_d(mlo + (m + 1) * step),
                 fr = mel_to_hz_d(mlo + (m + 2) * step);
    const double f = (double)k * (double)sr / (double)n_fft;
    if (f >= fl && f <= fc && fc != fl) w = (f - fl) / (fc - fl);
    else if (f > fc && f <= fr && fr != fc) w = (fr - f) / (fr - fc);
    w *= 2.0 / (fr - fl);
  }
  const float wf = (float)w;
  fb32[i] = wf;
  fbb[i] = f2bf(wf);
  fbh[i] = f2h(wf);
}

// Chroma filterbank (Chroma.swift:158-275, librosa filters.chroma with
// ctroct 5, octwidth 2, L2 column norm, base C), one thread per column k of
// the fbs-wide table: FFT bin k > 0 at n_chroma log2(k sr / n_fft / (A440/16))
// fractional chroma bins, DC 1.5 octaves below bin 1, width
// max(next - this, 1) (1 for bin n_fft - 1), Gaussian bumps of the wrapped
// distance, column L2 norm, octave weight, rows rolled by -3 n_chroma / 12.
// Rows >= n_chroma and columns > n_fft/2 are zero. FP64 throughout.
__global__ void k_tab_chroma(float* fb32, uint16_t* fbb, uint16_t* fbh, int fbs, int rows, int sr, int n_fft,
                             int n_chroma, double tuning) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= fbs) return;
  const int F = n_fft / 2 + 1;
  double w[kMaxMelBlocks * kMelBlock];
  const int nc = min(n_chroma, kMaxMelBlocks * kMelBlock);
  for (int c = 0; c < nc; ++c) w[c] = 0.0;
  if (k < F && nc > 0) {
    const double a16 = 440.0 * exp2(tuning / n_chroma) / 16.0;
    auto frq = [&](int b) -> double {
      if (b == 0) return n_chroma * log2(1.0 * sr / n_fft / a16) - 1.5 * n_chroma;
      return n_chroma * log2((double)b * sr / n_fft / a16);
    };
    const double fk = frq(k);
    const double width = (k < n_fft - 1) ? fmax(frq(k + 1) - fk, 1.0) : 1.0;
    const double half = rint(n_chroma / 2.0);
    double norm = 0.0;
    for (int c = 0; c < nc; ++c) {
      double d = fmod(fk - c + half + 10.0 * n_chroma, (double)n_chroma);
      if (d < 0.0) d += n_chroma;
      d -= half;
      const double s = 2.0 * d / width;
      w[c] = exp(-0.5 * s * s);
      norm += w[c] * w[c];
    }
    norm = sqrt(norm);
    const double oct = fk / n_chroma - 5.0;
    const double ow = exp(-0.5 * (oct / 2.0) * (oct / 2.0));
    for (int c = 0; c < nc; ++c) w[c] = (norm > 0.0 ? w[c] / norm : w[c]) * ow;
  }
  const int shift = 3 * (n_chroma / 12);
  for (int r = 0; r < rows; ++r) {
    double v = 0.0;
    if (r < nc) v = w[(r + shift) % nc];   // row r holds chroma (r + shift) mod n_chroma
    const int64_t i = (int64_t)r * fbs + k;
    const float vf = (float)v;
    fb32[i] = vf;
    fbb[i] = f2bf(vf);
    fbh[i] = f2h(vf);
  }
}

// per block of 8 mel rows: first nonzero bin and chunk counts (16 and 32 bins)
__global__ void k_tab_bands(const float* fb32, int fbs, int F, int nmh, int2* band32, int2* band16) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nmh) return;
  int lo = F, hi = -1;
  for (int r = 0; r < kMelBlock; ++r) {
    const float* row = fb32 + (int64_t)(kMelBlock * b + r) * fbs;
    for (int k = 0; k < F; ++k)
      if (row[k] != 0.0f) { lo = min(lo, k); hi = max(hi, k); }
  }
  if (hi < 0) { band32[b] = make_int2(0, 0); band16[b] = make_int2(0, 0); return; }
  const int s32 = lo & ~3, s16 = lo & ~7;
  band32[b] = make_int2(s32, (hi + 1 - s32 + 15) / 16);
  band16[b] = make_int2(s16, (hi + 1 - s16 + 31) / 32);
}

// assign[pc][nwi][w * (32 / nw) + i]: block lists of mel_block_assign for
// pc 0 = 16-bin chunks (FP32), 1 = 32-bin chunks (16-bit) and nwi 0 = 8
// waves, 1 = 16 waves.
// cache_src[pc][..]: sources of the LDS filterbank cache (cache_sources).
__global__ void k_tab_assign(const int2* band32, const int2* band16, int nmh, int fbs, int* assign, int* cache_src) {
  if (blockIdx.x != 0 || threadIdx.x >= 4) return;
  const int pc = threadIdx.x >> 1, nwi = threadIdx.x & 1;
  mel_block_assign(pc ? band16 : band32, nmh, nwi ? 16 : 8, assign + (pc * 2 + nwi) * kMaxMelBlocks);
  if (nwi == 0 && nmh <= kMaxMelBlocks)
    (void)cache_sources(pc ? band16 : band32, nmh, fbs, pc ? 32 : 16, cache_src + pc * kMaxCacheChunks);
}

__device__ inline double dct_val(int c, int m, int n_mfcc, int n_mels) {
  if (c >= n_mfcc || m >= n_mels) return 0.0;
  const double norm = (c == 0) ? sqrt(1.0 / n_mels) : sqrt(2.0 / n_mels);
  return cospi((double)c * (2.0 * m + 1.0) / (2.0 * n_mels)) * norm;
}

// dct32 [ndt][nmh][4][64]: lane l, step j -> D[16ct + l%16][8b + 4(l/16) + j]
// for l/16 < 2 and zero for l/16 >= 2 (the MFMA K slots that a block of 8
// mels leaves empty); dct16 [ndt][nmh][64][4]: lane l, elem j -> same entry
// (hi / lo split)
__global__ void k_tab_dct(float* dct32, uint16_t* bhi, uint16_t* blo, uint16_t* hhi, uint16_t* hlo, float* dense,
                          int ndt, int nmh, int n_mfcc, int n_mels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)ndt * nmh * 256;
  if (i < n) {
    const int lane = (int)(i % 64);
    const int j = (int)((i / 64) % 4);
    const int b = (int)((i / 256) % nmh);
    const int ct = (int)(i / (256 * (int64_t)nmh));
    const int g = lane >> 4;
    const int c = 16 * ct + (lane & 15), m = kMelBlock * b + 4 * g + j;
    const double d = (g < 2) ? dct_val(c, m, n_mfcc, n_mels) : 0.0;
    dct32[i] = (float)d;
    const int64_t o = (((int64_t)ct * nmh + b) * 64 + lane) * 4 + j;
    const uint16_t h = f2bf((float)d);
    bhi[o] = h;
    blo[o] = f2bf((float)(d - (double)bf2f(h)));
    const uint16_t hh = f2h((float)d);
    hhi[o] = hh;
    hlo[o] = f2h((float)(d - (double)h2f(hh)));
  }
  if (i < (int64_t)n_mfcc * n_mels) {
    const int c = (int)(i / n_mels), m = (int)(i % n_mels);
    dense[i] = (float)dct_val(c, m, n_mfcc, n_mels);
  }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t gen_tables(const TableKey& key, const DeviceTables& t, hipStream_t st) {
  const int M = key.n_fft / 2;
  const int F = key.n_fft / 2 + 1;
  hipLaunchKernelGGL(k_tab_window, dim3(nblk(key.n_fft, 256)), dim3(256), 0, st, t.window, key.n_fft, key.win);
  if (M >= 1)
    hipLaunchKernelGGL(k_tab_twiddle, dim3(nblk(M, 256)), dim3(256), 0, st, t.tw, M, t.twN, key.n_fft);
  {  // per-pass tables of the fused FFT: radix-16 passes then the tail radix
    const int log2m = (M >= 1) ? __builtin_ctz(M) : 0;
    const int np16 = log2m / 4, tail = 1 << (log2m % 4);
    const int npass = np16 + (tail > 1 ? 1 : 0);
    int off = 0;
    for (int p = 1; p < npass; ++p) {
      const int R = (p <= np16 - 1) ? 16 : tail, NS = 1 << (4 * p);
      hipLaunchKernelGGL(k_tab_twiddle_pass, dim3(nblk((int64_t)(R - 1) * NS, 256)), dim3(256), 0, st, t.twp + off,
                         NS, R);
      off += (R - 1) * NS;
    }
  }
  if (t.mft) hipLaunchKernelGGL(k_tab_mfft, dim3(nblk(kMfTabBytes / 16, 256)), dim3(256), 0, st, t.window, t.mft);
  const int rows = t.nmb * 16;
  if (rows > 0) {
    if (key.fb_kind == FB_CHROMA)
      hipLaunchKernelGGL(k_tab_chroma, dim3(nblk(t.fbs, 64)), dim3(64), 0, st, t.fb32, t.fb16, t.fbh, t.fbs, rows,
                         key.sr, key.n_fft, key.n_mels, (double)key.tuning);
    else
      hipLaunchKernelGGL(k_tab_mel, dim3(nblk((int64_t)rows * t.fbs, 256)), dim3(256), 0, st, t.fb32, t.fb16,
                         t.fbh, t.fbs, rows, key.sr, key.n_fft, key.n_mels, (double)key.fmin, (double)key.fmax);
    hipLaunchKernelGGL(k_tab_bands, dim3(nblk(t.nmh, 64)), dim3(64), 0, st, t.fb32, t.fbs, F, t.nmh, t.band32,
                       t.band16);
    hipLaunchKernelGGL(k_tab_assign, dim3(1), dim3(64), 0, st, t.band32, t.band16, t.nmh, t.fbs, t.assign,
                       t.cache_src);
  }
  const int64_t nd = std::max<int64_t>((int64_t)t.ndt * t.nmh * 256, (int64_t)key.n_mfcc * key.n_mels);
  if (nd > 0)
    hipLaunchKernelGGL(k_tab_dct, dim3(nblk(nd, 256)), dim3(256), 0, st, t.dct32, t.dctb_hi, t.dctb_lo, t.dcth_hi,
                       t.dcth_lo, t.dct_dense, t.ndt, t.nmh, key.n_mfcc, key.n_mels);
  return hipGetLastError();
}

// ---- fused kernel dispatch ----

static int g_tile_waves;

template <int LOG2M, int MODE, int PREC, int NW, int FPW, typename TIn, bool PIPE = false>
static hipError_t launch_tile_t(const TileArgs& a, int grid, size_t lds, hipStream_t st) {
  auto fn = k_tile<LOG2M, MODE, PREC, NW, FPW, TIn, PIPE>;
  // set once per instantiation (thread-safe static init); grid == 0 only
  // prepares the attribute, so stream capture never sees the first call
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit_bytes());
  if (attr != hipSuccess) return attr;
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(NW * 64), lds, st, a);
  return hipGetLastError();
}

// (NW, FPW): 16 x 1 up to n_fft 2048; 8 x 1 (8-frame tiles) at every size;
// 8 x 2 (16-frame tiles) from n_fft 1024 to 4096
template <int LOG2M, int MODE, int PREC, typename TIn>
static hipError_t launch_tile_s(const TileArgs& a, const TileLayout& L, int grid, hipStream_t st) {
  if constexpr (LOG2M <= 10) {
    if (L.shape == SHAPE_W16) return launch_tile_t<LOG2M, MODE, PREC, 16, 1, TIn>(a, grid, L.lds_bytes, st);
  }
  if constexpr (mfft_kernel(false, PREC, LOG2M) && MODE == OUT_MFCC) {
    if (L.shape == SHAPE_W16P) return launch_tile_t<LOG2M, MODE, PREC, 16, 1, TIn, true>(a, grid, L.lds_bytes, st);
  }
  if constexpr (LOG2M >= 9 && LOG2M <= 11) {
    if (L.shape == SHAPE_W8F2) return launch_tile_t<LOG2M, MODE, PREC, 8, 2, TIn>(a, grid, L.lds_bytes, st);
  }
  if (L.shape != SHAPE_W8) return hipErrorInvalidValue;
  return launch_tile_t<LOG2M, MODE, PREC, 8, 1, TIn>(a, grid, L.lds_bytes, st);
}

void set_tile_waves(int nw) { g_tile_waves = (nw == 8 || nw == 16) ? nw : 0; }
int tile_waves() { return g_tile_waves; }

template <int LOG2M>
static hipError_t launch_tile_m(const TileArgs& a, int prec, const TileLayout& L, int grid, hipStream_t st) {
  const bool h = (a.in_dtype == IN_F16);
  switch (a.out_kind) {
    case OUT_STFT_MAG: return launch_tile_s<LOG2M, OUT_STFT_MAG, PREC_F32, float>(a, L, grid, st);
    case OUT_STFT_POWER: return launch_tile_s<LOG2M, OUT_STFT_POWER, PREC_F32, float>(a, L, grid, st);
    case OUT_STFT_COMPLEX: return launch_tile_s<LOG2M, OUT_STFT_COMPLEX, PREC_F32, float>(a, L, grid, st);
    case OUT_MEL:
      if (prec == PREC_BF16)
        return h ? launch_tile_s<LOG2M, OUT_MEL, PREC_BF16, _Float16>(a, L, grid, st)
                 : launch_tile_s<LOG2M, OUT_MEL, PREC_BF16, float>(a, L, grid, st);
      if (prec == PREC_F16)
        return h ? launch_tile_s<LOG2M, OUT_MEL, PREC_F16, _Float16>(a, L, grid, st)
                 : launch_tile_s<LOG2M, OUT_MEL, PREC_F16, float>(a, L, grid, st);
      return h ? launch_tile_s<LOG2M, OUT_MEL, PREC_F32, _Float16>(a, L, grid, st)
               : launch_tile_s<LOG2M, OUT_MEL, PREC_F32, float>(a, L, grid, st);
    case OUT_MFCC:
      if (prec == PREC_BF16)
        return h ? launch_tile_s<LOG2M, OUT_MFCC, PREC_BF16, _Float16>(a, L, grid, st)
                 : launch_tile_s<LOG2M, OUT_MFCC, PREC_BF16, float>(a, L, grid, st);
      if (prec == PREC_F16)
        return h ? launch_tile_s<LOG2M, OUT_MFCC, PREC_F16, _Float16>(a, L, grid, st)
                 : launch_tile_s<LOG2M, OUT_MFCC, PREC_F16, float>(a, L, grid, st);
      return h ? launch_tile_s<LOG2M, OUT_MFCC, PREC_F32, _Float16>(a, L, grid, st)
               : launch_tile_s<LOG2M, OUT_MFCC, PREC_F32, float>(a, L, grid, st);
  }
  return hipErrorInvalidValue;
}

static size_t g_lds_limit = 64 * 1024;
void set_lds_limit_bytes(size_t n) { g_lds_limit = n; }
size_t lds_limit_bytes() { return g_lds_limit; }

int tile_shapes(int n_fft, int* out) {
  if (n_fft < 256 || n_fft > 4096 || (n_fft & (n_fft - 1))) return 0;
  int n = 0;
  if (n_fft <= 2048) out[n++] = SHAPE_W16;
  out[n++] = SHAPE_W8;
  if (n_fft >= 1024) out[n++] = SHAPE_W8F2;
  if (n_fft == 2048) out[n++] = SHAPE_W16P;   // its layout exists for the 16-bit MFCC kernels
  return n;
}

bool tile_layout(int n_fft, int out_kind, int prec, int n_mels, int n_mfcc, TileLayout* L, int shape) {
  if (n_fft < 256 || n_fft > 4096 || (n_fft & (n_fft - 1))) return false;
  const int M = n_fft / 2, F = M + 1;
  if (shape == SHAPE_DEFAULT) {
    if (M > 1024) {
      // n_fft 4096: 16-frame tiles of two frames per wave (faster for every
      // output kind and precision), 8-frame tiles where their LDS does not fit
      if (tile_layout(n_fft, out_kind, prec, n_mels, n_mfcc, L, SHAPE_W8F2)) return true;
      shape = SHAPE_W8;
    } else if (g_tile_waves != 0) {   // METALMOM_AMD_TILE_WAVES: 8 or 16 waves, one frame each
      shape = (g_tile_waves == 8) ? SHAPE_W8 : SHAPE_W16;
    } else {
      // n_fft 1024 / 2048: the STFT magnitude / power in 16-frame tiles of
      // two frames per wave, the 16-bit MFCC at 2048 pipelined (each measured
      // faster than 16 single-frame waves); W16 for everything else
      if ((M == 512 || M == 1024) && out_kind <= OUT_STFT_POWER &&
          tile_layout(n_fft, out_kind, prec, n_mels, n_mfcc, L, SHAPE_W8F2))
        return true;
      if (M == 1024 && out_kind == OUT_MFCC && prec != PREC_F32 &&
          tile_layout(n_fft, out_kind, prec, n_mels, n_mfcc, L, SHAPE_W16P))
        return true;
      shape = SHAPE_W16;
    }
  }
  if (shape != SHAPE_W16 && shape != SHAPE_W8 && shape != SHAPE_W8F2 && shape != SHAPE_W16P) return false;
  if (shape == SHAPE_W16 && M > 1024) return false;   // 16 waves need n_fft <= 2048
  if (shape == SHAPE_W8F2 && M != 512 && M != 1024 && M != 2048) return false;   // register-split FFT sizes
  const bool spec = out_kind <= OUT_STFT_COMPLEX;
  if (!spec) {
    if (n_mels < 1 || n_mels > kMaxMelBlocks * kMelBlock) return false;   // larger: generic path
    if (out_kind == OUT_MFCC && (n_mfcc < 1 || n_mfcc > 128)) return false;
  }
  const bool pipe = (shape == SHAPE_W16P);
  const int nw = (shape == SHAPE_W16 || pipe) ? 16 : 8, fpw = (shape == SHAPE_W8F2) ? 2 : 1;
  // FFT scratch: float2 exchanges for the small sizes, half-size float
  // exchanges for the register-split sizes (n_fft >= 1024), none for the
  // matrix-core transform (registers only)
  const bool mf = mfft_kernel(spec, prec, __builtin_ctz(M));
  if (pipe && !(mf && out_kind == OUT_MFCC)) return false;
  const int scr = mf ? 0 : (M + M / 16) * ((M >= 512) ? 4 : 8);
  int esz, S;
  if (spec) {
    esz = (out_kind == OUT_STFT_COMPLEX) ? 8 : 4;
    S = F + 1;
  } else if (prec == PREC_F32) {
    esz = 4;
    const int need = ((F + 31) / 32) * 32 + 64;
    S = need + ((4 - need % 64) + 64) % 64;          // S = 4 (mod 64) dwords
  } else {
    esz = 2;
    const int need = ((F + 31) / 32) * 32 + 64;
    S = need + ((8 - need % 128) + 128) % 128;       // S/2 = 4 (mod 64) dwords
  }
  // region: the FFT scratch with the wave's last staging row over it, then
  // the rows of its other frames
  const int first = std::max((scr + 15) / 16 * 16, (S * esz + 15) / 16 * 16);
  int rb = (fpw > 1) ? first + (fpw - 1) * S * esz : std::max(S * esz, scr);
  rb = (rb + 15) / 16 * 16;
  // frame exponents and per-wave extrema (one set per staging buffer), phase-2 flags
  const size_t extras = (16 * 4 + 2 * nw * 4) * (pipe ? 2 : 1) + nw * 4;
  size_t total = (size_t)nw * rb + extras;
  if (total > lds_limit_bytes()) return false;
  int tab_off = 0;
  if (mf) {
    // the matrix-core transform's table in LDS (any shape; it holds the window)
    tab_off = (int)((total + 15) / 16 * 16);
    total = (size_t)tab_off + (size_t)kMfTabBytes;
    if (total > lds_limit_bytes()) return false;
  } else if (!spec && prec != PREC_F32 && nw == 16 && (M == 512 || M == 1024)) {
    // FFT tables in LDS (the kernel's TABL case): window, per-pass twiddles, W_N^k
    tab_off = (int)((total + 15) / 16 * 16);
    total = (size_t)tab_off + (size_t)n_fft * 4 + (size_t)M * 8 * 2;
    if (total > lds_limit_bytes()) return false;
  }
  int part_off = 0;
  if (out_kind == OUT_MFCC) {
    const int ndt = (n_mfcc + 15) / 16;
    const int nmh = (n_mels + kMelBlock - 1) / kMelBlock;   // one partial per mel block
    const size_t part = (size_t)ndt * nmh * 256 * 4;
    if (total + part + 16 <= lds_limit_bytes()) {
      // DCT partials in their own area: two barriers per tile
      part_off = (int)(((total + 15) / 16) * 16);
      total = part_off + part;
    } else if (part > (size_t)nw * rb) {
      return false;   // partials would not even fit over the staging regions
    }
    if (pipe) {
      // the second staging regions and partial area
      if (part_off == 0) return false;
      L->stage2_off = (int)((total + 15) / 16 * 16);
      total = (size_t)L->stage2_off + (size_t)nw * rb;
      L->part2_off = (int)((total + 15) / 16 * 16);
      total = (size_t)L->part2_off + part;
      if (total > lds_limit_bytes()) return false;
    }
  }
  if (!pipe) {
    L->stage2_off = 0;
    L->part2_off = 0;
  }
  L->nw = nw;
  L->fpw = fpw;
  L->row_stride = S;
  L->region_bytes = rb;
  L->lds_bytes = total;
  L->part_off = part_off;
  L->tab_off = tab_off;
  L->tile_frames = nw * fpw;
  L->shape = shape;
  // the two-frame shape exists to run two workgroups per CU
  L->occ = (shape == SHAPE_W16 || pipe) ? 1 : (int)std::max<size_t>(1, std::min<size_t>(2, lds_limit_bytes() / total));
  return true;
}

hipError_t launch_tile(const TileArgs& a, int prec, const TileLayout& L, int grid, hipStream_t st) {
  const int log2m = __builtin_ctz(a.n_fft) - 1;
#ifdef MM_DEV_ONLY_LOG2M  // development builds: one FFT size only
  if (log2m == MM_DEV_ONLY_LOG2M) return launch_tile_m<MM_DEV_ONLY_LOG2M>(a, prec, L, grid, st);
  return hipErrorInvalidValue;
#else
  switch (log2m) {
    case 7: return launch_tile_m<7>(a, prec, L, grid, st);
    case 8: return launch_tile_m<8>(a, prec, L, grid, st);
    case 9: return launch_tile_m<9>(a, prec, L, grid, st);
    case 10: return launch_tile_m<10>(a, prec, L, grid, st);
    case 11: return launch_tile_m<11>(a, prec, L, grid, st);
  }
  return hipErrorInvalidValue;
#endif
}

hipError_t launch_fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(std::min<unsigned>(nblk(n, 256), 1024)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_generic_fft(const float* x, int64_t len, int n_fft, int hop, int center, int pad_mode,
                              const float* window, const float2* tw, const float2* twN, int T, float2* spec,
                              hipStream_t st) {
  if (T <= 0) return hipSuccess;
  const size_t lds = (n_fft >= 2) ? (size_t)n_fft * 8 : 16;
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void*>(k_generic_fft), hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)lds_limit_bytes());
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(k_generic_fft, dim3(T), dim3(256), lds, st, x, len, n_fft, hop, center, pad_mode, window, tw,
                     twN, T, spec);
  return hipGetLastError();
}

hipError_t launch_generic_spec_out(const float2* spec, int F, int T, int kind, void* out, hipStream_t st) {
  const int64_t n = (int64_t)F * T;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_generic_spec_out, dim3(nblk(n, 256)), dim3(256), 0, st, spec, F, T, kind, out);
  return hipGetLastError();
}

hipError_t launch_generic_mel(const float2* spec, int F, int T, const float* fb, int fbs, int n_mels, float power,
                              float* mel, hipStream_t st) {
  const int64_t n = (int64_t)n_mels * T;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_generic_mel, dim3(nblk(n, 256)), dim3(256), 0, st, spec, F, T, fb, fbs, n_mels, power, mel);
  return hipGetLastError();
}

hipError_t launch_db(const float* in, int64_t n, float mult, float ref, float amin, float* out, unsigned* max_key,
                     hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_db, dim3(std::min<unsigned>(nblk(n, 256), 1024)), dim3(256), 0, st, in, n, mult, ref, amin,
                     out, max_key);   // 256 threads = the 4 waves k_db reduces over
  return hipGetLastError();
}

hipError_t launch_db_clamp(float* x, int64_t n, const unsigned* max_key, float top_db, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_db_clamp, dim3(std::min<unsigned>(nblk(n, 256), 2048)), dim3(256), 0, st, x, n, max_key,
                     top_db);
  return hipGetLastError();
}

hipError_t launch_generic_dct(const float* logmel, int n_mels, int T, const float* dct, int n_mfcc, float* out,
                              hipStream_t st) {
  const int64_t n = (int64_t)n_mfcc * T;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_generic_dct, dim3(nblk(n, 256)), dim3(256), 0, st, logmel, n_mels, T, dct, n_mfcc, out);
  return hipGetLastError();
}

}  // namespace mm
