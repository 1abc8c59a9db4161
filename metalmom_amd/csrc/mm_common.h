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

constexpr int kTileFrames = 16;   // frames per tile = MFMA N dimension

// Device-resident tables for one (sr, n_fft, win, n_mels, fmin, fmax, n_mfcc) key.
struct DeviceTables {
  float* window = nullptr;      // [n_fft]
  float2* tw = nullptr;         // [M] W_M^t, M = n_fft/2
  float2* twN = nullptr;        // [M] W_N^k, N = n_fft
  float2* twp = nullptr;        // per-pass Stockham twiddles of the fused FFT (< M entries)
  float* fb32 = nullptr;        // [nmb*16][fbs] mel weights, zero padded
  uint16_t* fb16 = nullptr;     // same in bf16 or f16 (see fb16_kind)
  uint16_t* fbh = nullptr;      // f16 copy
  int2* band32 = nullptr;       // [nmb] (k_start, n_chunks of 16 bins)
  int2* band16 = nullptr;       // [nmb] (k_start, n_chunks of 32 bins)
  float* dct32 = nullptr;       // [ndt][nmb][4][64] permuted DCT-II rows (A operand)
  uint16_t* dctb_hi = nullptr;  // [ndt][nmb][64][4] bf16 hi part
  uint16_t* dctb_lo = nullptr;  // bf16 lo part
  uint16_t* dcth_hi = nullptr;  // f16 hi part
  uint16_t* dcth_lo = nullptr;  // f16 lo part
  float* dct_dense = nullptr;   // [n_mfcc][n_mels] plain basis (generic path)
  int fbs = 0;                  // fb row stride (elements)
  int nmb = 0;                  // mel blocks of 16
  int ndt = 0;                  // DCT row tiles of 16
};

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
  int n_tiles;
  int tiles_per_wg;            // contiguous chunk of tiles per persistent workgroup
  // framing
  int n_fft, hop, center, pad_mode, in_dtype;
  // tables
  const float* window;
  const float2* tw;
  const float2* twN;
  const float2* twp;
  const float* fb32;
  const uint16_t* fb16;
  const int2* band;
  int fbs, nmb, n_mels;
  const float* dct32;
  const uint16_t* dct_hi;
  const uint16_t* dct_lo;
  int ndt, n_mfcc;
  // value transforms
  float power, ref, amin, top_db;
  float post_scale;            // 2^-power: the fused kernel contracts (2|X|)^power
  int use_top_db;
  int phase;                   // 1: unclamped pass, 2: recompute tiles whose clamp is active
  unsigned* clip_max_key;      // per clip, ordered-int encoding of the max dB
  float* tile_min;             // per tile min dB (phase 1 writes, phase 2 reads)
  // output
  void* out;
  int out_kind;
  int lds_row_stride;          // staging row stride in elements
  int lds_region_bytes;        // per-wave LDS region
  int lds_part_off;            // MFCC: LDS offset of the DCT partials (0: over the regions)
};

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
