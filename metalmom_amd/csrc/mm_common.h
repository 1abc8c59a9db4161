// Shared definitions between the HIP kernels and the C++ host engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mm {

// What a fused tile launch writes.
enum OutKind : int {
  OUT_STFT_MAG = 0,      // |X|           [F, T] float32   (mm_stft)
  OUT_STFT_POWER = 1,    // |X|^2         [F, T] float32
  OUT_STFT_COMPLEX = 2,  // X             [F, T] complex64 interleaved
  OUT_MEL = 3,           // FB @ |X|^p    [n_mels, T] float32 (mm_mel_spectrogram)
  OUT_MFCC = 4,          // DCT(dB(mel))  [n_mfcc, T] float32 (mm_mfcc)
};

// Arithmetic of the mel and DCT contractions.
enum Precision : int { PREC_F32 = 0, PREC_BF16 = 1, PREC_F16 = 2 };

// Sample type of the packed input signal.
enum InDtype : int { IN_F32 = 0, IN_F16 = 1 };

enum PadMode : int { PAD_CONSTANT = 0, PAD_REFLECT = 1 };

// Filterbank of the OUT_MEL / OUT_MFCC contraction: Slaney mel
// (FilterBank.swift:23-73) or the Gaussian chroma filterbank
// (Chroma.swift:158-275; n_mels = n_chroma rows).
enum FbKind : int { FB_MEL = 0, FB_CHROMA = 1 };

// Per-frame reductions of the magnitude spectrogram (SpectralDescriptors.swift).
enum DescKind : int { DESC_CENTROID = 0, DESC_BANDWIDTH = 1, DESC_ROLLOFF = 2, DESC_FLATNESS = 3 };

// Interleaved little-endian PCM sample formats (the values of MM_PCM_*).
enum PcmFormat : int { PCM_U8 = 1, PCM_S16 = 2, PCM_S24 = 3, PCM_S32 = 4, PCM_F32 = 5, PCM_F64 = 6 };
__host__ __device__ inline int pcm_bytes(int fmt) {
  return fmt == PCM_U8 ? 1 : fmt == PCM_S16 ? 2 : fmt == PCM_S24 ? 3 : fmt == PCM_F64 ? 8 : 4;
}
inline bool pcm_valid(int fmt) { return fmt >= PCM_U8 && fmt <= PCM_F64; }

constexpr int kTileFrames = 16;   // frames per tile = MFMA N dimension
// LDS copy of the banded filterbank: per chunk 8 rows of 64 bytes of weights,
// each row padded to 80 bytes so the 16-byte operand reads of lanes 0..7
// (rows 0..7) fall on distinct banks
constexpr int kFbRowBytes = 80;
constexpr int kFbChunkBytes = 8 * kFbRowBytes;
constexpr int kMelBlock = 8;      // mel rows per filterbank block of the fused contraction
constexpr int kMaxMelBlocks = 32; // n_mels <= 256 on the fused path

// Per tile: where its frames start in the packed signal and which of them can
// be fetched by the plain interior path (used to prefetch the next tile).
struct TilePf {
  long long base;   // packed-signal index of frame f0's first sample (may be < clip start)
  unsigned mask;    // bit w: frame f0 + w exists, lies inside its clip and starts at an even offset
  int pad;
};

// Device-resident tables for one (sr, n_fft, win, n_mels, fmin, fmax, n_mfcc) key.
struct DeviceTables {
  float* window = nullptr;      // [n_fft]
  float2* tw = nullptr;         // [M] W_M^t, M = n_fft/2
  float2* twN = nullptr;        // [M] W_N^k, N = n_fft
  float2* twp = nullptr;        // per-pass Stockham twiddles of the fused FFT (< M entries)
  unsigned char* mft = nullptr; // matrix-core FFT table (n_fft 2048 only; mfft2048)
  float* fb32 = nullptr;        // [nmb*16][fbs] mel weights, zero padded
  uint16_t* fb16 = nullptr;     // same in bf16 or f16 (see fb16_kind)
  uint16_t* fbh = nullptr;      // f16 copy
  int2* band32 = nullptr;       // [nmh] (k_start, n_chunks of 16 bins) per block of 8 mels
  int2* band16 = nullptr;       // [nmh] (k_start, n_chunks of 32 bins)
  int* assign = nullptr;        // [2 chunk widths][2 wave counts (8, 16)][32]: blocks per wave
  int* cache_src = nullptr;     // [2 chunk widths][kMaxCacheChunks]: filterbank cache sources
  float* dct32 = nullptr;       // [ndt][nmh][4][64] permuted DCT-II rows (A operand)
  uint16_t* dctb_hi = nullptr;  // [ndt][nmh][64][4] bf16 hi part
  uint16_t* dctb_lo = nullptr;  // bf16 lo part
  uint16_t* dcth_hi = nullptr;  // f16 hi part
  uint16_t* dcth_lo = nullptr;  // f16 lo part
  float* dct_dense = nullptr;   // [n_mfcc][n_mels] plain basis (generic path)
  int fbs = 0;                  // fb row stride (elements)
  int nmb = 0;                  // filterbank rows / 16 (rows are zero padded to 16 nmb)
  int nmh = 0;                  // mel blocks of 8 (the contraction's unit of work)
  int ndt = 0;                  // DCT row tiles of 16
};

// Table of the matrix-core FFT of the fused kernel (mm_fft_mfma.h), built
// once per configuration (it holds the analysis window) and copied into LDS
// by each workgroup:
//   [0, 8 KiB)        window pairs of each lane's input points, [pair 8][lane 64] x 4 floats
//   [8 KiB, 16 KiB)   constant operand, [parity p 2][part t 2][hi, lo][lane 64] x 8 FP16
//   [16 KiB, 24 KiB)  stage twiddles W_1024^(n1 k2), [p 2][mt 2][pair 2][lane 64] x 2 float2
//   [24 KiB, 28 KiB)  split twiddles W_2048^k, [p 2][pair 2][lane 64] x 2 float2
constexpr int kMfWinBytes = 8 * 64 * 16;
constexpr int kMfOpBytes = 2 * 2 * 2 * 64 * 16;
constexpr int kMfTw1Bytes = 8 * 64 * 16;
constexpr int kMfTw2Bytes = 4 * 64 * 16;
constexpr int kMfOpOff = kMfWinBytes;
constexpr int kMfTw1Off = kMfOpOff + kMfOpBytes;
constexpr int kMfTw2Off = kMfTw1Off + kMfTw1Bytes;
constexpr int kMfTabBytes = kMfTw2Off + kMfTw2Bytes;

// Whether the fused kernel of an output kind (spec: STFT outputs), operand
// precision and FFT size 2^(log2m+1) runs the matrix-core FFT: the 16-bit mel
// and MFCC kernels at n_fft 2048 (the FP32 kernels keep the FP32 FFT).
__host__ __device__ constexpr bool mfft_kernel(bool spec, int prec, int log2m) {
  return !spec && prec != PREC_F32 && log2m == 10;
}

// Everything one fused-tile launch needs. Passed by value as a kernel argument.
struct TileArgs {
  // packed signals
  const void* signal;
  const int64_t* clip_off;     // first sample of each clip in `signal`
  const int64_t* clip_len;     // samples per clip
  const int32_t* clip_frames;  // frames per clip
  const int64_t* out_off;      // first output element of each clip
  const int32_t* tile_clip;    // clip of each tile
  const int32_t* tile_f0;      // first frame of each tile
  const TilePf* tile_pf;       // prefetch record of each tile
  int n_tiles;
  int tiles_per_wg;            // contiguous chunk of tiles per persistent workgroup
  // framing
  int n_fft, hop, center, pad_mode, in_dtype;
  // tables
  const float* window;
  const float2* tw;
  const float2* twN;
  const float2* twp;
  const unsigned char* mft;    // matrix-core FFT table (16-bit mel / MFCC at n_fft 2048)
  const float* fb32;
  const uint16_t* fb16;
  const int2* band;            // per block of 8 mels
  const int* assign;           // [waves][64 / waves] block ids of each wave (-1: none)
  int fbs, nmh, n_mels;
  const float* dct32;
  const uint16_t* dct_hi;
  const uint16_t* dct_lo;
  int ndt, n_mfcc;
  // value transforms
  float power, ref, amin, top_db;
  float db_off;                // 10 log10(ref): dB = 10 log10(2) log2(max(mel, amin)) - db_off
  float post_scale;            // 2^-power: the fused kernel contracts (2|X|)^power
  int use_top_db;
  int phase;                   // 1: unclamped pass, 2: recompute tiles whose clamp is active
  unsigned* clip_max_key;      // per clip, ordered-int encoding of the max dB
  // phase 1, when set: the clip maxima of the next step, which workgroup 0
  // sets to -inf (a plan alternates two sets, so its steps need no fill
  // launch; nullptr: the caller clears clip_max_key before phase 1)
  unsigned* clip_max_clear;
  int n_clips;
  float* tile_min;             // per tile min dB (phase 1 writes, phase 2 reads)
  // output
  void* out;
  int out_kind;
  int lds_row_stride;          // staging row stride in elements
  int lds_region_bytes;        // per-wave LDS region
  int lds_part_off;            // MFCC: LDS offset of the DCT partials (0: over the regions)
  int lds_stage2_off;          // pipelined shape: second staging regions (odd processed tiles)
  int lds_part2_off;           // pipelined shape: second DCT partial area
  int tab_lds_off;             // LDS copies of window and twiddles (16-bit mel / MFCC kernels)
  int fb_lds_off;              // LDS copy of the banded filterbank (0: read it from global memory)
  int dct_lds_off;             // LDS copy of the DCT operand tables (0: global memory)
  const int* cache_src;        // [cache_chunks] sources of the filterbank cache (cache_sources)
  int cache_chunks;
  long long* phase_cycles;     // instrumented builds only: [grid][waves][9] cycle sums
};

// A wave's mel block slot, packed in 32 bits: bit 31 set for a used slot,
// bits 0-4 the block, 5-14 its first bin / 4, 15-22 its chunks, 23-30 the
// index of its first chunk in the LDS filterbank cache (blocks in order,
// chunks back to back; meaningful when the total is at most kMaxCacheChunks).
constexpr int kMaxCacheChunks = 255;
__host__ __device__ inline int slot_pack(int block, int first_bin, int chunks, int cache_first) {
  return (int)(0x80000000u | (unsigned)block | ((unsigned)(first_bin >> 2) << 5) | ((unsigned)chunks << 15) |
               ((unsigned)(cache_first & 255) << 23));
}
__host__ __device__ inline bool slot_used(int v) { return v < 0; }
__host__ __device__ inline int slot_block(int v) { return v & 31; }
__host__ __device__ inline int slot_first_bin(int v) { return ((v >> 5) & 1023) << 2; }
__host__ __device__ inline int slot_chunks(int v) { return (v >> 15) & 255; }
__host__ __device__ inline int slot_cache_first(int v) { return (int)(((unsigned)v >> 23) & 255u); }

// Mel blocks per wavefront of the fused contraction, longest processing time
// first: blocks in order of decreasing cost (chunks + 1; highest index first on
// ties) each go to the least loaded wave that still has a free slot (lowest
// index on ties). The first min(nw, nmh) blocks land on waves 0 .. nw-1, so
// the waves holding blocks are always 0 .. min(nw, nmh)-1. Writes
// out[w * (kMaxMelBlocks / nw) + i] = slot_pack(...) of the block, 0 for an
// empty slot (all slots when nmh is 0 or above kMaxMelBlocks). band[b] =
// (first bin, chunks) of block b; first bins are multiples of 4.
__host__ __device__ inline void mel_block_assign(const int2* band, int nmh, int nw, int* out) {
  const int slots = kMaxMelBlocks / nw;
  for (int i = 0; i < kMaxMelBlocks; ++i) out[i] = 0;
  if (nmh < 1 || nmh > kMaxMelBlocks || nw < 1 || nw > 16) return;
  int cfirst[kMaxMelBlocks];   // first cache chunk of each block
  int acc = 0;
  for (int b = 0; b < nmh; ++b) {
    cfirst[b] = acc;
    acc += band[b].y;
  }
  int load[16], cnt[16];
  for (int w = 0; w < 16; ++w) { load[w] = 0; cnt[w] = 0; }
  bool taken[kMaxMelBlocks];
  for (int b = 0; b < kMaxMelBlocks; ++b) taken[b] = false;
  for (int n = 0; n < nmh; ++n) {
    int best = -1;
    for (int b = 0; b < nmh; ++b)
      if (!taken[b] && (best < 0 || band[b].y >= band[best].y)) best = b;
    taken[best] = true;
    int wb = -1;   // nw * slots >= nmh, so a wave with a free slot exists
    for (int w = 0; w < nw; ++w)
      if (cnt[w] < slots && (wb < 0 || load[w] < load[wb])) wb = w;
    out[wb * slots + cnt[wb]] = slot_pack(best, band[best].x, band[best].y, cfirst[best]);
    cnt[wb] += 1;
    load[wb] += band[best].y + 1;
  }
}

// Source of each LDS filterbank cache chunk: the element offset of its row 0
// (mel 8 b of block b) at its first bin, chunks of `ch` bins, blocks in
// order. Returns the number of chunks (0 when more than kMaxCacheChunks).
__host__ __device__ inline int cache_sources(const int2* band, int nmh, int fbs, int ch, int* src) {
  int n = 0;
  for (int b = 0; b < nmh; ++b) n += band[b].y;
  if (n > kMaxCacheChunks) return 0;
  int i = 0;
  for (int b = 0; b < nmh; ++b)
    for (int c = 0; c < band[b].y; ++c) src[i++] = kMelBlock * b * fbs + band[b].x + ch * c;
  return n;
}

// Ordered encoding of a float so that unsigned comparison matches float order.
__host__ __device__ inline unsigned float_to_key(float x) {
  unsigned u;
  __builtin_memcpy(&u, &x, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float key_to_float(unsigned k) {
  unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

}  // namespace mm
