// Launchers for the HIP kernels in mm_kernels.hip (host-callable, internal).
#pragma once
#include "mm_common.h"

namespace mm {

// Key of one set of device-resident tables.
struct TableKey {
  int sr, n_fft, win, n_mels, n_mfcc;
  float fmin, fmax;
  int fb_kind = FB_MEL;   // filterbank of the n_mels rows
  float tuning = 0.0f;    // chroma: deviation from A440 in fractional chroma bins
  bool operator==(const TableKey& o) const {
    return sr == o.sr && n_fft == o.n_fft && win == o.win && n_mels == o.n_mels && n_mfcc == o.n_mfcc &&
           fmin == o.fmin && fmax == o.fmax && fb_kind == o.fb_kind && tuning == o.tuning;
  }
};

// LDS layout of the fused tile kernel for one configuration.
struct TileLayout {
  int nw;            // wavefronts per workgroup
  int fpw;           // frames per wavefront
  int tile_frames;   // frames per tile (nw * fpw)
  int row_stride;    // staging row stride, elements
  int region_bytes;  // LDS bytes per wavefront region
  size_t lds_bytes;  // dynamic LDS per workgroup
  int part_off;      // LDS offset of the DCT partial area (0: reuse the staging regions)
  int tab_off;       // LDS offset of the FFT table copies (16-bit mel / MFCC, n_fft 1024, 2048)
  int shape;         // TileShape this layout was made for (never SHAPE_DEFAULT)
  int stage2_off;    // SHAPE_W16P: LDS offset of the second staging regions (0: single-buffered)
  int part2_off;     // SHAPE_W16P: LDS offset of the second DCT partial area
  int occ;           // workgroups per CU the layout is sized for (LDS copies of tables only within this)
};

// Workgroup shapes of the fused kernel, chosen per batch by the autotuner.
enum TileShape : int {
  SHAPE_DEFAULT = 0,   // 16 waves x 1 frame up to n_fft 2048, 8 x 1 at 4096 (or set_tile_waves)
  SHAPE_W16 = 1,       // 16 waves, one frame each: 16-frame tiles (n_fft <= 2048)
  SHAPE_W8 = 2,        // 8 waves, one frame each: 8-frame tiles
  SHAPE_W8F2 = 3,      // 8 waves, two frames each in turn: 16-frame tiles, two workgroups per CU
                       // (n_fft 1024 and 2048), so one workgroup's contraction and barriers
                       // overlap the other's FFTs
  SHAPE_W16P = 4,      // 16 waves, one frame each, pipelined: staging rows and DCT partials
                       // double-buffered, one barrier per tile, so waves that finish a tile's
                       // contraction start the next FFT while others still contract
                       // (16-bit MFCC at n_fft 2048: the matrix-core FFT needs no LDS scratch)
};
// Shapes the fused kernel offers for this FFT size (SHAPE_W16 first when it
// exists); returns their number (at most 4).
int tile_shapes(int n_fft, int* out);

void set_lds_limit_bytes(size_t n);
// Workgroup shape of the fused kernel for n_fft <= 2048: 16 or 8 waves (one
// frame each); 0 restores the default (16).
void set_tile_waves(int nw);
int tile_waves();
size_t lds_limit_bytes();

hipError_t gen_tables(const TableKey& key, const DeviceTables& t, hipStream_t st);

// Returns false when the fused kernel does not cover this configuration
// (then the generic kernels are used).
bool tile_layout(int n_fft, int out_kind, int prec, int n_mels, int n_mfcc, TileLayout* L, int shape = SHAPE_DEFAULT);
hipError_t launch_tile(const TileArgs& a, int prec, const TileLayout& L, int grid, hipStream_t st);

hipError_t launch_fill_u32(unsigned* p, int64_t n, unsigned v, hipStream_t st);

hipError_t launch_generic_fft(const float* x, int64_t len, int n_fft, int hop, int center, int pad_mode,
                              const float* window, const float2* tw, const float2* twN, int T, float2* spec,
                              hipStream_t st);
hipError_t launch_generic_spec_out(const float2* spec, int F, int T, int kind, void* out, hipStream_t st);
hipError_t launch_generic_mel(const float2* spec, int F, int T, const float* fb, int fbs, int n_mels, float power,
                              float* mel, hipStream_t st);
hipError_t launch_db(const float* in, int64_t n, float mult, float ref, float amin, float* out, unsigned* max_key,
                     hipStream_t st);
hipError_t launch_db_clamp(float* x, int64_t n, const unsigned* max_key, float top_db, hipStream_t st);
hipError_t launch_generic_dct(const float* logmel, int n_mels, int T, const float* dct, int n_mfcc, float* out,
                              hipStream_t st);

// Onset strength from the dB mel spectrogram db [n_bands, T] (before the
// top_db floor) and its maximum key: out [n_out] (aggregate) or
// [n_bands, n_out]; output frame o holds the flux at i = o - pad.
hipError_t launch_onset(const float* db, int n_bands, int T, const unsigned* max_key, float top_db, int lag, int pad,
                        int n_out, int aggregate, float* out, hipStream_t st);

// Spectral descriptor (DescKind) of every frame of S [F, T] -> out [T];
// param: bandwidth p, rolloff roll_percent.
hipError_t launch_descriptor(const float* S, int F, int T, int kind, float srf, float nfftf, float param,
                             float* out, hipStream_t st);
// Spectral contrast bands: nb = n_bands + 1 bands, band b = bins
// [edge[b], edge[b+1]), q[b] values averaged at each end.
constexpr int kMaxContrastBands = 63;
struct ContrastBands {
  int nb;
  int edge[kMaxContrastBands + 1];
  int q[kMaxContrastBands];
};
// out [nb, T] (row-major) from the magnitude spectrogram S [F, T]; d_cb: the
// band table in device memory.
hipError_t launch_contrast(const float* S, int T, int nb, const ContrastBands* d_cb, float* out, hipStream_t st);

// Fused descriptor of one clip (FFT + reduction per wave, no spectrogram in
// HBM) for n_fft 1024 / 2048 / 4096: desc_fused_waves gives the waves per
// workgroup (tiles of that many frames; 0 = no fused kernel); out[T].
int desc_fused_waves(int n_fft);
hipError_t launch_desc_fused(const TileArgs& a, int kind, float srf, float nfftf, float param, float* out, int grid,
                             hipStream_t st);

// ---- PCM ingest ----
// out[frames] = interleaved PCM (PcmFormat) -> float32: the mean of the
// channels (mix) or the first channel.
hipError_t launch_pcm_to_mono(const void* raw, int64_t frames, int channels, int fmt, int mix, float* out,
                              int n_cu, hipStream_t st);

// ---- sample-rate conversion ----
// FP32 filter table (sinc x Kaiser, 64 zero crossings, 512 points per
// crossing); resample_table_len() floats.
hipError_t launch_tab_resample(float* table, hipStream_t st);
int resample_table_len();
// y[n_out] = x[n] converted from source_sr to target_sr (rates differ, both
// > 0); grid: persistent workgroups (one per CU).
hipError_t launch_resample(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                           const float* table, int grid, hipStream_t st);
// Polyphase form: true when this rate pair's weight rows (rows 0 .. phases/2
// of stride floats; the others are their mirror images) fit in LDS;
// launch_poly_table writes those rows ((phases / 2 + 1) * stride floats)
// from the filter table, launch_resample_poly converts with them.
bool resample_poly_shape(int source_sr, int target_sr, int* phases, int* taps, int* stride);
hipError_t launch_poly_table(int source_sr, int target_sr, const float* table, float* wtab, hipStream_t st);
// wsum1: the sum of the weight row when there is one phase (its sign decides
// whether interior outputs are 0, as the per-output weight sum would).
hipError_t launch_resample_poly(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                const float* wtab, float wsum1, int grid, hipStream_t st);
// One phase with source = 2 target: two outputs per lane from float4 reads
// of the staged window (wtab, wsum1 as for launch_resample_poly).
hipError_t launch_resample_dec2(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                const float* wtab, float wsum1, int grid, hipStream_t st);
// Phase-major form for rate pairs with 2 .. 4096 phases whose transposed
// input window (64 outputs of each phase per wave) fits in LDS: the full
// table of phases rows of stride floats, then one row sum per phase
// (phases * (stride + 1) floats, launch_phase_table), read by scalar loads.
struct PhaseShape {
  int phases, step, support, taps, stride, groups, ld;
};
bool resample_phase_shape(int source_sr, int target_sr, PhaseShape* sh);
hipError_t launch_phase_table(int source_sr, int target_sr, const float* table, float* wfull, hipStream_t st);
hipError_t launch_resample_phase(const float* x, int64_t n, float* y, int64_t n_out, int source_sr, int target_sr,
                                 const float* wfull, int grid, hipStream_t st);

// ---- inverse STFT ----
struct IstftArgs {
  const float2* X;       // [n_fft/2 + 1][ld] complex, row-major (the reference's interleaved floats at ld = T)
  float* out;            // [out_len]
  const float* window;   // [n_fft] synthesis window
  const float2* tw;      // W_M^k, k < M (generic path)
  const float2* twN;     // W_N^k, k < M
  const float2* twp;     // per-pass twiddles of the register-split FFT
  int n_fft, hop, T;
  int ld;                // row stride of X in complex elements (>= T)
  int halo;              // fused: frames before a tile that touch its samples, ceil(n_fft / hop) - 1
  int carry;             // fused: carried partial sums per tile, n_fft - hop
  int carry_stride;      // fused: floats per carry buffer (carry rounded up to 4)
  int tile0;             // fused: index of the first tile (tile t = frames t*nw .. t*nw + nw - 1)
  int n_tiles;           // fused: tiles of this launch
  int tiles_per_wg;      // fused: consecutive tiles per workgroup (persistent grid)
  int region_bytes;      // fused: LDS bytes per wave
  int64_t total;         // overlap-add buffer length, n_fft + (T - 1) hop
  int64_t start;         // first buffer sample of the output (n_fft/2 with centre, else 0)
  int64_t out_len;       // output samples (zeros past the buffer)
};
struct IstftPlan {
  int nw, halo, carry, carry_stride, region_bytes;
  size_t lds_bytes;
};
// Fused tile kernel plan: false when n_fft is not 1024 / 2048 / 4096 or the
// halo (ceil(n_fft / hop) - 1 frames) exceeds half a tile.
bool istft_fused_plan(int n_fft, int hop, IstftPlan* p);
hipError_t launch_istft_tile(const IstftArgs& a, int nw, int grid, size_t lds, hipStream_t st);
// Per-frame inverse FFTs into frames[T][n_fft], then the overlap-add.
hipError_t launch_istft_generic(const IstftArgs& a, float* frames, hipStream_t st);

}  // namespace mm
