// Host engine: device singleton, table cache, per-context stream/workspace
// lanes, single-clip and batched execution. Internal to libmetalmom.so.
#pragma once
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/metalmom.h"
#include "mm_common.h"
#include "mm_kernels.h"

namespace mm {

// Status codes (values of the public header).
enum Status : int32_t { OK = 0, ERR_INVALID = -1, ERR_NO_GPU = -2, ERR_ALLOC = -3, ERR_INTERNAL = -4 };

// Full description of one spectral request.
struct SpecRequest {
  int out_kind = OUT_MFCC;
  int sr = 22050;
  int n_fft = 2048;
  int hop = 512;
  int win = 2048;
  int center = 1;
  int pad_mode = PAD_CONSTANT;
  float power = 2.0f;
  int n_mels = 128;
  float fmin = 0.0f;
  float fmax = 11025.0f;
  int n_mfcc = 13;
  float ref = 1.0f;
  float amin = 1e-10f;
  float top_db = 80.0f;   // <= 0: no clamp
  int precision = PREC_F32;
  int in_dtype = IN_F32;
  int fb_kind = FB_MEL;   // FB_CHROMA: n_mels chroma rows (fmin / fmax unused)
  float tuning = 0.0f;    // chroma tuning, fractional chroma bins
};

int64_t frame_count(int64_t n, int n_fft, int hop, int center);
// rows of the output matrix for a request (bins, mels or coefficients)
int64_t out_rows(const SpecRequest& r);
// float32 elements per output entry (2 for complex)
int out_elem_floats(const SpecRequest& r);

struct Tables;  // device-resident tables (cached, process lifetime)

// Grow-only device / pinned host buffers.
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t bytes);
  void release();
};
struct HostBuf {
  void* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t bytes);
  void release();
};

// One stream plus its workspaces. Pooled: contexts borrow one on first use.
//
// The *_device entry points enqueue work that reads the lane's buffers on the
// caller's stream and return before it runs. `busy` is recorded after such
// work; any later use of the lane from another stream (or from the host
// paths on `stream`, or the lane's return to the pool) first waits for it,
// so buffers are neither rewritten nor reallocated under a running kernel.
// Calls on the stream that recorded it are ordered by the stream itself.
struct Lane {
  hipStream_t stream = nullptr;
  DevBuf d_in, d_out, d_meta, d_scratch, d_scratch2, d_small;
  HostBuf h_in, h_out, h_meta, h_small;
  std::string meta_key;   // work list now in d_meta for device-resident single-clip calls ("" = none)
  std::string bands_key;  // spectral contrast band table now in d_small (at kBandsOff), "" = none
  hipEvent_t busy = nullptr;          // recorded after the last device-entry work
  hipStream_t busy_stream = nullptr;  // the stream it was recorded on
  bool busy_set = false;
  // before using the lane's buffers for work on `st`
  int32_t enter(hipStream_t st);
  // after enqueuing device-entry work on `st` that uses the lane's buffers
  int32_t leave(hipStream_t st);
  // waits for every recorded device-entry work
  int32_t settle();
  ~Lane();
};

class Engine {
 public:
  static Engine* get();          // nullptr when no GPU is usable
  static int32_t status();       // OK or ERR_NO_GPU

  int num_cus() const { return n_cu_; }
  const std::string& arch() const { return arch_; }
  const std::string& name() const { return name_; }

  Lane* acquire_lane();
  void release_lane(Lane* l);

  // Tables for a request (generated on first use, then shared).
  int32_t tables_for(const SpecRequest& r, hipStream_t st, const Tables** out);
  // The request's tables if they are already resident, else nullptr (never allocates).
  const Tables* tables_resident(const SpecRequest& r);
  // Filter table of the sample-rate converter (generated on first use).
  int32_t resample_table(hipStream_t st, const float** out);
  // Polyphase weight rows of a rate pair (built on first use; nullptr when
  // they do not fit the polyphase kernel); *wsum1: the sum of the row when
  // there is one phase.
  int32_t resample_weights(int source_sr, int target_sr, hipStream_t st, const float** out, float* wsum1);
  // Full weight table and row sums of the phase-major kernel (built on first
  // use; nullptr when the rate pair does not fit it).
  int32_t resample_phase_weights(int source_sr, int target_sr, hipStream_t st, const float** out);

  // Persistent grid size: autotuned for this shape if known, else a default.
  int grid_for(const SpecRequest& r, int n_tiles, const TileLayout& L);
  // Records the autotuner's winner: its workgroup shape and, for that shape,
  // the workgroups per CU.
  void record_tuning(const SpecRequest& r, int shape, int wg_per_cu);
  int tuned_wg_per_cu(const SpecRequest& r, int shape);
  int tuned_shape(const SpecRequest& r);   // SHAPE_DEFAULT when not tuned
  std::string tuning_path() const;
  void load_tuning();
  void save_tuning();

 private:
  Engine() = default;
  bool init();
  int n_cu_ = 0;
  std::string arch_, name_;
  std::mutex mu_;
  std::vector<Lane*> free_lanes_;
  std::vector<std::unique_ptr<Tables>> tables_;
  DevBuf rs_table_;
  struct PolyEntry {
    int src, dst;
    bool full;   // phase-major table (all rows and their sums)
    float wsum1;
    DevBuf w;
  };
  std::vector<std::unique_ptr<PolyEntry>> poly_;
  std::vector<std::pair<std::string, int>> tune_;   // tuning file entries (key, value)
  bool tune_loaded_ = false;
};

std::string tune_key(const SpecRequest& r, const std::string& arch);

// Device-side batch description (shared by single-clip and batched paths).
struct BatchMeta {
  int n_clips = 0;
  int n_tiles = 0;
  int64_t total_in = 0;     // samples
  int64_t total_out = 0;    // float32 elements
  std::vector<int64_t> clip_off, clip_len, out_off;
  std::vector<int32_t> clip_frames, tile_clip, tile_f0;
  std::vector<TilePf> tile_pf;
};
void build_meta(const SpecRequest& r, const int64_t* lens, int n, int tile_frames, BatchMeta* m);

// Device execution of a whole batch whose signal and output already live in
// HBM. `d_meta` must hold the uploaded metadata (see meta_bytes/pack_meta).
size_t meta_bytes(const BatchMeta& m);
void pack_meta(const BatchMeta& m, void* host);
struct MetaPtrs {
  const int64_t* clip_off;
  const int64_t* clip_len;
  const int32_t* clip_frames;
  const int64_t* out_off;
  const int32_t* tile_clip;
  const int32_t* tile_f0;
  unsigned* clip_max_key;
  unsigned* clip_max_key2;   // the second set of a plan's alternating steps
  float* tile_min;
  const TilePf* tile_pf;
};
MetaPtrs meta_ptrs(const BatchMeta& m, void* dev);

// Workgroups of a fused launch over n_tiles tiles (grid_override > 0:
// workgroups requested, else the tuned or default grid) and the contiguous
// tiles each one processes.
int launch_grid(Engine* eng, const SpecRequest& r, int n_tiles, const TileLayout& L, int grid_override,
                int* tiles_per_wg);
// `shape`: the TileShape the metadata's tiles were built for.
int32_t enqueue_batch(Engine* eng, const SpecRequest& r, const Tables* t, const BatchMeta& m, const MetaPtrs& mp,
                      const void* d_signal, void* d_out, DevBuf* scratch, DevBuf* scratch2, int grid_override,
                      hipStream_t st, hipEvent_t ev_p1_begin = nullptr, hipEvent_t ev_p1_end = nullptr,
                      int shape = SHAPE_DEFAULT, unsigned* clear_next = nullptr);

// Device buffer for the phase timing of instrumented builds (null: off).
long long* phase_cycles_buffer();
void set_phase_cycles_buffer(long long* p);

// Sets the fused kernel's LDS attribute ahead of any stream capture.
int32_t prepare_batch(const SpecRequest& r, int shape = SHAPE_DEFAULT);

// Single clip through pinned staging on the lane's stream (synchronous).
// On success *out is malloc'd (caller frees) and rows/cols are set.
int32_t run_single(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, float** out, int64_t* rows,
                   int64_t* cols);

int32_t run_db(Lane* lane, const float* data, int64_t n, float mult, float ref, float amin, float top_db,
               float** out);

// MMSpecParams of the additive API -> request, with its sentinels (hop <= 0
// -> n_fft/4, win <= 0 -> n_fft, f_max <= 0 -> sr/2) and validation; 16-bit
// contractions and FP16 input only where the fused kernel covers them.
int32_t request_from_params(const MMSpecParams* p, SpecRequest* r);

// Onset strength of one clip (OnsetDetection.swift:23-128): r is the power-2
// mel request (win = n_fft); aggregate -> [1, frames], else [n_mels, frames].
// One zero column when the clip has at most `lag` frames.
struct OnsetShape {
  int T = 0;       // STFT frames
  int pad = 0;     // leading zeros
  int rows = 0;    // 1 (aggregate) or n_mels
  int cols = 0;    // output frames
};
OnsetShape onset_shape(const SpecRequest& r, int64_t len, int aggregate, int lag);
// Device signal d_x (len floats) -> d_res (rows x cols floats), enqueued on
// st; the mel spectrogram and the maximum key live in the lane's buffers and
// the clip's work list is uploaded once per (length, request).
int32_t enqueue_onset(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int aggregate, int lag,
                      float* d_res, hipStream_t st);
// Device signal d_x (len floats) -> descriptor per frame in d_res, enqueued on
// st; the magnitude spectrogram lives in the lane's buffers, the clip's work
// list is uploaded once per (length, request).
int32_t enqueue_descriptor(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int kind, float param,
                           float* d_res, hipStream_t st);
// Spectral descriptor of one clip: r is the magnitude STFT request; *out
// (malloc'd) gets one value per frame, *frames their count. Synchronous.
int32_t run_descriptor(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int kind, float param,
                       float** out, int64_t* frames);
// Host in, host out through the lane (synchronous); *out is malloc'd.
// Status of a failed HIP call; MM_HIP returns it from the enclosing function.
#define MM_HIP(expr)                                 \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_status(e_);     \
  } while (0)

inline int32_t hip_status(hipError_t e) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return ERR_ALLOC;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return ERR_NO_GPU;
  return ERR_INTERNAL;
}
inline int32_t hip_status_or_ok(hipError_t e) { return e == hipSuccess ? OK : hip_status(e); }

// The lane of a live mm_context (nullptr for NULL or a destroyed context).
Lane* ctx_lane(void* ctx);
// Interleaved PCM (PcmFormat) -> one channel (mean when mix, else channel 0)
// -> resampled from source_sr to target_sr, all on the GPU; malloc'd result.
int32_t run_pcm_resample(Lane* lane, const void* raw, int64_t frames, int channels, int fmt, int mix,
                         int source_sr, int target_sr, float** out, int64_t* n_out);
// Sample-rate conversion (Resample.swift:12-96): output length
// ceil(n * target / source) in double, n when the rates agree.
int64_t resample_length(int64_t n, int source_sr, int target_sr);
// Enqueues d_y[resample_length(n)] = d_x[n] converted on st (a device copy
// when the rates agree).
int32_t enqueue_resample(const float* d_x, int64_t n, int source_sr, int target_sr, float* d_y, hipStream_t st);
// Host clip -> malloc'd result of *n_out samples (nullptr when empty).
int32_t run_resample(Lane* lane, const float* x, int64_t n, int source_sr, int target_sr, float** out,
                     int64_t* n_out);
// Spectral contrast (SpectralDescriptors.swift:137-204) of one host clip:
// bands from contrast_bands, result [n_bands + 1, frames] (malloc'd).
bool contrast_bands(int sr, int n_fft, int n_freqs, int n_bands, float fmin, ContrastBands* cb);
int32_t run_contrast(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int n_bands, float fmin,
                     float** out, int64_t* frames);
// The same on device memory: d_res [n_bands + 1, frames], enqueued on st
// (the band table is uploaded, with a host synchronisation, once per
// configuration on a lane).
int32_t enqueue_contrast(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int n_bands, float fmin,
                         float* d_res, hipStream_t st);
int32_t run_onset(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int aggregate, int lag,
                  float** out, int64_t* rows, int64_t* cols);

// ---- inverse STFT ----
// Output geometry of mm_istft (STFT.swift:703-704, 720-724, 832-858).
struct IstftShape {
  int n_fft = 0, hop = 0, T = 0;
  bool pow2 = false;     // false: the result is total zeros, untrimmed
  int64_t total = 0;     // overlap-add buffer, n_fft + (T - 1) hop
  int64_t start = 0;     // buffer sample of output sample 0
  int64_t out_len = 0;   // samples of the result
};
// ERR_INVALID for shapes the library does not take (n_freqs or n_frames < 1,
// hop < 0, a power-of-two n_fft above 16384, a buffer of 2^31 samples or more).
int32_t istft_shape(int n_freqs, int n_frames, int hop, int center, int64_t length, IstftShape* sh);
// d_X: [n_freqs][ld] complex on the device (ld >= n_frames; a multiple of 16
// lets the fused kernel read whole cache lines); d_out: sh.out_len floats.
// Scratch (generic path) comes from `scratch`. Enqueued on st.
int32_t enqueue_istft(const IstftShape& sh, int sr, int win, const float* d_X, int ld, float* d_out,
                      DevBuf* scratch, hipStream_t st);
// Host in, host out through the lane (synchronous); *out is malloc'd
// (nullptr when the result is empty).
int32_t run_istft(Lane* lane, const IstftShape& sh, int sr, int win, const float* X, float** out);

}  // namespace mm
