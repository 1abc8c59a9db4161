// C ABI of libmetalmom.so (declared in include/metalmom.h).
//
// Validation, sentinels, ownership and empty-result conventions follow the
// reference bridge, Sources/MetalMomBridge/Bridge.swift:
//   mm_init/mm_destroy registry      :16-40
//   mm_stft                          :44-140
//   mm_istft                         :144-221
//   mm_onset_strength                :1319-1358
//   mm_chroma_stft                   :494-545
//   mm_spectral_{centroid,bandwidth,contrast,rolloff,flatness}  :621-813
//   mm_amplitude_to_db               :225-263
//   mm_power_to_db                   :267-306
//   mm_mel_spectrogram               :310-400
//   mm_mfcc                          :402-493
//   mm_resample                      :1087-1104
//   mm_buffer_free                   :3695-3711
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <mutex>
#include <unordered_set>
#include <vector>

#include "../../include/metalmom.h"
#include "mm_engine.h"

using namespace mm;

namespace {

struct Ctx {
  Lane* lane = nullptr;
  Lane* acquire() {
    if (!lane) {
      Engine* e = Engine::get();
      if (!e) return nullptr;
      lane = e->acquire_lane();
    }
    return lane;
  }
  // host-buffer paths (work on the lane's own stream, synchronous)
  Lane* get() { return for_stream(acquire() ? lane->stream : nullptr); }
  // device-buffer paths: work on the caller's stream `st`
  Lane* for_stream(hipStream_t st) {
    Lane* l = acquire();
    if (!l || l->enter(st) != OK) return nullptr;
    return l;
  }
};

// Records the end of a device-entry call's work on the lane (see Lane).
int32_t device_call_done(Lane* lane, hipStream_t st, int32_t s) {
  if (s != OK) return s;
  return lane->leave(st);
}

std::mutex g_reg_mu;
std::unordered_set<void*>& registry() {
  static std::unordered_set<void*> r;
  return r;
}

bool is_live(void* p) {
  std::lock_guard<std::mutex> g(g_reg_mu);
  return registry().count(p) != 0;
}

bool pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

void fill_shape(MMBuffer* out, const std::vector<int64_t>& shape, int dtype, int64_t count, float* data) {
  out->data = data;
  for (int i = 0; i < 8; ++i) out->shape[i] = 0;
  for (size_t i = 0; i < shape.size() && i < 8; ++i) out->shape[i] = shape[i];
  out->ndim = (int32_t)shape.size();
  out->dtype = dtype;
  out->count = count;
}

// Runs one clip and fills `out` ([rows, frames]; empty results keep the shape
// with data = NULL and count = 0, Bridge.swift:95-113).
int32_t run_to_buffer(void* ctx, const SpecRequest& r, const float* sig, int64_t len, MMBuffer* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  const int64_t rows = out_rows(r);
  const int64_t T = frame_count(len, r.n_fft, r.hop, r.center);
  const int dtype = (r.out_kind == OUT_STFT_COMPLEX) ? 1 : 0;
  if (rows * T == 0) {
    fill_shape(out, {rows, T}, dtype, 0, nullptr);
    return OK;
  }
  Lane* lane = c->get();
  if (!lane) return ERR_NO_GPU;
  float* data = nullptr;
  int64_t rr = 0, cc = 0;
  const int32_t s = run_single(lane, r, sig, len, &data, &rr, &cc);
  if (s != OK) return s;
  fill_shape(out, {rr, cc}, dtype, rr * cc, data);
  return OK;
}

// Spectral descriptor of one clip into a 1-D [frames] buffer (Bridge.swift:621-813):
// hop <= 0 -> n_fft/4, win <= 0 -> n_fft (STFT.compute defaults); a
// non-power-of-two n_fft has no frames (empty result).
int32_t descriptor_to_buffer(void* ctx, const float* sig, int64_t len, int32_t sr, int32_t n_fft, int32_t hop,
                             int32_t win, int32_t center, int kind, float param, MMBuffer* out) {
  if (!ctx || !sig || len <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (n_fft <= 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_STFT_MAG;
  r.sr = sr;
  r.n_fft = n_fft;
  r.hop = hop > 0 ? hop : n_fft / 4;
  r.win = win > 0 ? win : n_fft;
  r.center = center != 0;
  if (r.hop <= 0) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft) || frame_count(len, n_fft, r.hop, r.center) == 0) {
    fill_shape(out, {0}, 0, 0, nullptr);
    return OK;
  }
  Lane* lane = static_cast<Ctx*>(ctx)->get();
  if (!lane) return ERR_NO_GPU;
  float* data = nullptr;
  int64_t T = 0;
  const int32_t s = run_descriptor(lane, r, sig, len, kind, param, &data, &T);
  if (s != OK) return s;
  fill_shape(out, {T}, 0, T, data);
  return OK;
}

int32_t db_to_buffer(void* ctx, const float* data, int64_t count, float mult, float ref, float amin, float top_db,
                     MMBuffer* out) {
  Lane* lane = nullptr;
  if (ctx && is_live(ctx)) lane = static_cast<Ctx*>(ctx)->get();
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  bool borrowed = false;
  if (!lane) {  // the reference does not require a context here (Bridge.swift:277-279)
    lane = e->acquire_lane();
    borrowed = true;
    if (!lane) return ERR_INTERNAL;
  }
  float* res = nullptr;
  const int32_t s = run_db(lane, data, count, mult, ref, amin, top_db, &res);
  if (borrowed) e->release_lane(lane);
  if (s != OK) return s;
  fill_shape(out, {count}, 0, count, res);
  return OK;
}


}  // namespace

namespace mm {
Lane* ctx_lane(void* ctx) { return (ctx && is_live(ctx)) ? static_cast<Ctx*>(ctx)->get() : nullptr; }
}  // namespace mm

extern "C" {

mm_context mm_init(void) try {
  Ctx* c = new Ctx();
  std::lock_guard<std::mutex> g(g_reg_mu);
  registry().insert(c);
  return c;
} catch (...) {   // no exception leaves the C ABI
  return nullptr;
}

void mm_destroy(mm_context ctx) try {
  if (!ctx) return;
  {
    std::lock_guard<std::mutex> g(g_reg_mu);
    if (registry().erase(ctx) == 0) return;  // already destroyed: no-op
  }
  Ctx* c = static_cast<Ctx*>(ctx);
  if (c->lane) {
    (void)hipStreamSynchronize(c->lane->stream);
    Engine::get()->release_lane(c->lane);
  }
  delete c;
} catch (...) {   // no exception leaves the C ABI
}

int32_t mm_stft(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                const MMSTFTParams* params, MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !params || !out) return MM_ERR_INVALID_INPUT;
  if (!pow2(params->n_fft)) return MM_ERR_INVALID_INPUT;
  if (params->hop_length <= 0) return MM_ERR_INVALID_INPUT;  // the reference divides by it
  SpecRequest r;
  r.out_kind = OUT_STFT_MAG;
  r.sr = sample_rate;
  r.n_fft = params->n_fft;
  r.hop = params->hop_length;
  r.win = params->win_length;  // taken verbatim; <= 0 gives an all-zero window
  r.center = params->center != 0;
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_descriptor_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                                      int32_t sample_rate, int32_t n_fft, int32_t hop_length, int32_t win_length,
                                      int32_t center, int32_t kind, float param, float* d_out, int64_t* out_frames,
                                      void* stream) try {
  if (!ctx || !is_live(ctx) || !d_signal || signal_length <= 0 || !d_out) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft) || kind < MM_DESC_CENTROID || kind > MM_DESC_FLATNESS) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_STFT_MAG;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  if (out_frames) *out_frames = frame_count(signal_length, n_fft, r.hop, r.center);
  hipStream_t st = static_cast<hipStream_t>(stream);
  Lane* lane = static_cast<Ctx*>(ctx)->for_stream(st);
  if (!lane) return ERR_NO_GPU;
  return device_call_done(lane, st, enqueue_descriptor(lane, r, d_signal, signal_length, kind, param, d_out, st));
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_contrast_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                                    int32_t sample_rate, int32_t n_fft, int32_t hop_length, int32_t win_length,
                                    int32_t center, int32_t n_bands, float f_min, float* d_out, int64_t* out_frames,
                                    void* stream) try {
  if (!ctx || !is_live(ctx) || !d_signal || signal_length <= 0 || !d_out) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft) || sample_rate <= 0 || n_bands < 1 || n_bands + 1 > kMaxContrastBands) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_STFT_MAG;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  if (out_frames) *out_frames = frame_count(signal_length, n_fft, r.hop, r.center);
  hipStream_t st = static_cast<hipStream_t>(stream);
  Lane* lane = static_cast<Ctx*>(ctx)->for_stream(st);
  if (!lane) return ERR_NO_GPU;
  return device_call_done(lane, st, enqueue_contrast(lane, r, d_signal, signal_length, n_bands, f_min, d_out, st));
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_onset_strength_device(mm_context ctx, const float* d_signal, int64_t signal_length, int32_t sample_rate,
                                 int32_t n_fft, int32_t hop_length, int32_t n_mels, float f_min, float f_max,
                                 int32_t center, int32_t aggregate, float* d_out, int64_t* out_rows,
                                 int64_t* out_cols, void* stream) try {
  if (!ctx || !is_live(ctx) || !d_signal || signal_length <= 0 || !d_out) return MM_ERR_INVALID_INPUT;
  if (sample_rate <= 0 || n_mels < 1 || !pow2(n_fft)) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_MEL;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : 512;
  r.win = n_fft;
  r.center = center != 0;
  r.power = 2.0f;
  r.n_mels = n_mels;
  r.fmin = f_min;
  r.fmax = f_max > 0.0f ? f_max : (float)sample_rate / 2.0f;
  const OnsetShape o = onset_shape(r, signal_length, aggregate != 0, 1);
  if (out_rows) *out_rows = o.rows;
  if (out_cols) *out_cols = o.cols;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Lane* lane = static_cast<Ctx*>(ctx)->for_stream(st);
  if (!lane) return ERR_NO_GPU;
  return device_call_done(lane, st, enqueue_onset(lane, r, d_signal, signal_length, aggregate != 0, 1, d_out, st));
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int64_t mm_istft_length(int32_t n_freqs, int32_t n_frames, int32_t hop_length, int32_t center,
                        int64_t output_length) try {
  IstftShape sh;
  if (istft_shape(n_freqs, n_frames, hop_length, center, output_length, &sh) != OK) return -1;
  return sh.out_len;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_istft_device(mm_context ctx, const float* d_stft, int32_t ld_frames, int32_t n_freqs, int32_t n_frames,
                        int32_t sample_rate, int32_t hop_length, int32_t win_length, int32_t center,
                        int64_t output_length, float* d_out, void* stream) try {
  if (!ctx || !is_live(ctx) || !d_stft || !d_out) return MM_ERR_INVALID_INPUT;
  IstftShape sh;
  if (istft_shape(n_freqs, n_frames, hop_length, center, output_length, &sh) != OK) return MM_ERR_INVALID_INPUT;
  const int ld = (ld_frames > 0) ? ld_frames : n_frames;
  if (ld < n_frames) return MM_ERR_INVALID_INPUT;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Lane* lane = static_cast<Ctx*>(ctx)->for_stream(st);
  if (!lane) return ERR_NO_GPU;
  return device_call_done(lane, st, enqueue_istft(sh, sample_rate, win_length, d_stft, ld, d_out, &lane->d_scratch, st));
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_stft_complex(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                        const MMSTFTParams* params, MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !params || !out) return MM_ERR_INVALID_INPUT;
  if (!pow2(params->n_fft) || params->hop_length <= 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_STFT_COMPLEX;
  r.sr = sample_rate;
  r.n_fft = params->n_fft;
  r.hop = params->hop_length;
  r.win = params->win_length;
  r.center = params->center != 0;
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_istft(mm_context ctx, const float* stft_data, int64_t stft_count, int32_t n_freqs, int32_t n_frames,
                 int32_t sample_rate, int32_t hop_length, int32_t win_length, int32_t center, int64_t output_length,
                 MMBuffer* out) try {
  if (!ctx || !stft_data || stft_count <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (n_freqs < 1 || n_frames < 1 || stft_count != 2 * (int64_t)n_freqs * n_frames) return MM_ERR_INVALID_INPUT;
  IstftShape sh;
  if (istft_shape(n_freqs, n_frames, hop_length, center, output_length, &sh) != OK) return MM_ERR_INVALID_INPUT;
  if (sh.out_len == 0) {   // Bridge.swift:190-201: data NULL, shape all zero
    fill_shape(out, {0}, 0, 0, nullptr);
    return OK;
  }
  if (!sh.pow2) {   // STFT.swift:720-724: the zero buffer, nothing computed
    float* z = static_cast<float*>(std::calloc((size_t)sh.out_len, sizeof(float)));
    if (!z) return MM_ERR_ALLOC_FAILED;
    fill_shape(out, {sh.out_len}, 0, sh.out_len, z);
    return OK;
  }
  Lane* lane = static_cast<Ctx*>(ctx)->get();
  if (!lane) return ERR_NO_GPU;
  float* data = nullptr;
  const int32_t s = run_istft(lane, sh, sample_rate, win_length, stft_data, &data);
  if (s != OK) return s;
  fill_shape(out, {sh.out_len}, 0, sh.out_len, data);
  return OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_onset_strength(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                          int32_t n_fft, int32_t hop_length, int32_t n_mels, float f_min, float f_max, int32_t center,
                          int32_t aggregate, MMBuffer* out) try {
  // Bridge.swift:1334-1336 checks the signal, its length and out, not the context
  if (!signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (sample_rate <= 0 || n_mels < 1) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_MEL;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : 512;   // OnsetDetection.swift:38 (not n_fft/4)
  r.win = n_fft;
  r.center = center != 0;
  r.power = 2.0f;
  r.n_mels = n_mels;
  r.fmin = f_min;
  r.fmax = f_max > 0.0f ? f_max : (float)sample_rate / 2.0f;
  const int agg = aggregate != 0;
  // no STFT frames (a non-power-of-two n_fft) or a single frame: one zero
  // column (OnsetDetection.swift:63-71), nothing computed
  if (!pow2(n_fft) || frame_count(signal_length, n_fft, r.hop, r.center) <= 1) {
    const int64_t rows = agg ? 1 : n_mels;
    float* z = static_cast<float*>(std::calloc((size_t)rows, sizeof(float)));
    if (!z) return MM_ERR_ALLOC_FAILED;
    fill_shape(out, {rows, 1}, 0, rows, z);
    return OK;
  }
  Lane* lane = nullptr;
  if (ctx && is_live(ctx)) lane = static_cast<Ctx*>(ctx)->get();
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  bool borrowed = false;
  if (!lane) {
    lane = e->acquire_lane();
    borrowed = true;
    if (!lane) return ERR_INTERNAL;
  }
  float* data = nullptr;
  int64_t rows = 0, cols = 0;
  const int32_t s = run_onset(lane, r, signal_data, signal_length, agg, 1, &data, &rows, &cols);
  if (borrowed) e->release_lane(lane);
  if (s != OK) return s;
  fill_shape(out, {rows, cols}, 0, rows * cols, data);
  return OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_resample(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t source_sr,
                    int32_t target_sr, MMBuffer* out) try {
  // Bridge.swift:1095-1097 checks the signal, its length and out, not the context
  if (!signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (source_sr <= 0 || target_sr <= 0) return MM_ERR_INVALID_INPUT;
  if (source_sr == target_sr) {   // Resample.swift:13: the signal itself
    float* c = static_cast<float*>(std::malloc((size_t)signal_length * sizeof(float)));
    if (!c) return MM_ERR_ALLOC_FAILED;
    std::memcpy(c, signal_data, (size_t)signal_length * sizeof(float));
    fill_shape(out, {signal_length}, 0, signal_length, c);
    return OK;
  }
  Lane* lane = nullptr;
  if (ctx && is_live(ctx)) lane = static_cast<Ctx*>(ctx)->get();
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  bool borrowed = false;
  if (!lane) {
    lane = e->acquire_lane();
    borrowed = true;
    if (!lane) return ERR_INTERNAL;
  }
  float* data = nullptr;
  int64_t n = 0;
  const int32_t s = run_resample(lane, signal_data, signal_length, source_sr, target_sr, &data, &n);
  if (borrowed) e->release_lane(lane);
  if (s != OK) return s;
  fill_shape(out, {n}, 0, n, data);
  return OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int64_t mm_resample_length(int64_t signal_length, int32_t source_sr, int32_t target_sr) try {
  if (signal_length < 0 || source_sr <= 0 || target_sr <= 0) return -1;
  return resample_length(signal_length, source_sr, target_sr);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_resample_device(mm_context ctx, const float* d_signal, int64_t signal_length, int32_t source_sr,
                           int32_t target_sr, float* d_out, void* stream) try {
  (void)ctx;
  if (!d_signal || signal_length <= 0 || !d_out) return MM_ERR_INVALID_INPUT;
  if (source_sr <= 0 || target_sr <= 0) return MM_ERR_INVALID_INPUT;
  if (!Engine::get()) return ERR_NO_GPU;
  return enqueue_resample(d_signal, signal_length, source_sr, target_sr, d_out, static_cast<hipStream_t>(stream));
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_amplitude_to_db(mm_context ctx, const float* data, int64_t count, float ref, float amin, float top_db,
                           MMBuffer* out) try {
  if (!data || count <= 0 || !out) return MM_ERR_INVALID_INPUT;
  return db_to_buffer(ctx, data, count, 20.0f, ref, amin, top_db, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_power_to_db(mm_context ctx, const float* data, int64_t count, float ref, float amin, float top_db,
                       MMBuffer* out) try {
  if (!data || count <= 0 || !out) return MM_ERR_INVALID_INPUT;
  return db_to_buffer(ctx, data, count, 10.0f, ref, amin, top_db, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_mel_spectrogram(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                           int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, float power,
                           int32_t n_mels, float f_min, float f_max, MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (n_fft <= 0 || n_mels < 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_MEL;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  r.power = power;
  r.n_mels = n_mels;
  r.fmin = f_min;
  r.fmax = f_max > 0.0f ? f_max : (float)sample_rate / 2.0f;
  if (r.hop <= 0) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft)) {  // the reference core returns an empty [n_mels, 0] result
    fill_shape(out, {n_mels, 0}, 0, 0, nullptr);
    return MM_OK;
  }
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_chroma_stft(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                       int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t n_chroma, int32_t center,
                       MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (n_fft <= 0 || n_chroma < 1 || n_chroma > 256 || sample_rate <= 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_MEL;   // the banded contraction with the chroma filterbank
  r.fb_kind = FB_CHROMA;
  r.tuning = 0.0f;        // Bridge.swift:526-534 passes no tuning (Chroma.stft default 0)
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;   // Bridge.swift:521-524
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  r.power = 2.0f;
  r.n_mels = n_chroma;
  if (r.hop <= 0) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft)) {   // the STFT gives no frames: [n_chroma, 0]
    fill_shape(out, {n_chroma, 0}, 0, 0, nullptr);
    return MM_OK;
  }
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_centroid(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                             int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, MMBuffer* out) try {
  return descriptor_to_buffer(ctx, signal_data, signal_length, sample_rate, n_fft, hop_length, win_length, center,
                              DESC_CENTROID, 0.0f, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_bandwidth(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                              int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, float p,
                              MMBuffer* out) try {
  return descriptor_to_buffer(ctx, signal_data, signal_length, sample_rate, n_fft, hop_length, win_length, center,
                              DESC_BANDWIDTH, p, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_contrast(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                             int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, int32_t n_bands,
                             float f_min, MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;   // Bridge.swift:711-716
  // n_bands < 1 traps in the reference (the range 1...n_bands); sr <= 0 has no bin width
  if (n_fft <= 0 || sample_rate <= 0 || n_bands < 1 || n_bands + 1 > kMaxContrastBands) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_STFT_MAG;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  if (r.hop <= 0) return MM_ERR_INVALID_INPUT;
  const int64_t nb = n_bands + 1;
  if (!pow2(n_fft) || frame_count(signal_length, n_fft, r.hop, r.center) == 0) {   // no STFT frames
    fill_shape(out, {nb, 0}, 0, 0, nullptr);
    return OK;
  }
  Lane* lane = static_cast<Ctx*>(ctx)->get();
  if (!lane) return ERR_NO_GPU;
  float* data = nullptr;
  int64_t T = 0;
  const int32_t s = run_contrast(lane, r, signal_data, signal_length, n_bands, f_min, &data, &T);
  if (s != OK) return s;
  fill_shape(out, {nb, T}, 0, nb * T, data);
  return OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_rolloff(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                            int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, float roll_percent,
                            MMBuffer* out) try {
  return descriptor_to_buffer(ctx, signal_data, signal_length, sample_rate, n_fft, hop_length, win_length, center,
                              DESC_ROLLOFF, roll_percent, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_flatness(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate,
                             int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t center, MMBuffer* out) try {
  return descriptor_to_buffer(ctx, signal_data, signal_length, sample_rate, n_fft, hop_length, win_length, center,
                              DESC_FLATNESS, 0.0f, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_mfcc(mm_context ctx, const float* signal_data, int64_t signal_length, int32_t sample_rate, int32_t n_mfcc,
                int32_t n_fft, int32_t hop_length, int32_t win_length, int32_t n_mels, float f_min, float f_max,
                int32_t center, MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  if (n_fft <= 0 || n_mels < 0 || n_mfcc < 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  r.out_kind = OUT_MFCC;
  r.sr = sample_rate;
  r.n_fft = n_fft;
  r.hop = hop_length > 0 ? hop_length : n_fft / 4;
  r.win = win_length > 0 ? win_length : n_fft;
  r.center = center != 0;
  r.power = 2.0f;
  r.n_mels = n_mels;
  r.n_mfcc = n_mfcc;
  r.fmin = f_min;
  r.fmax = f_max > 0.0f ? f_max : (float)sample_rate / 2.0f;
  r.ref = 1.0f;
  r.amin = 1e-10f;
  r.top_db = 80.0f;
  if (r.hop <= 0) return MM_ERR_INVALID_INPUT;
  if (!pow2(n_fft)) {
    fill_shape(out, {std::min(n_mfcc, n_mels), 0}, 0, 0, nullptr);
    return MM_OK;
  }
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

void mm_buffer_free(MMBuffer* buf) try {
  if (!buf) return;
  if (buf->data) std::free(buf->data);
  buf->data = nullptr;
  buf->ndim = 0;
  buf->dtype = 0;
  buf->count = 0;
  for (int i = 0; i < 8; ++i) buf->shape[i] = 0;
} catch (...) {   // no exception leaves the C ABI
}

int32_t mm_spectral(mm_context ctx, const MMSpecParams* params, const float* signal_data, int64_t signal_length,
                    MMBuffer* out) try {
  if (!ctx || !signal_data || signal_length <= 0 || !out) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  int32_t s = request_from_params(params, &r);
  if (s != OK) return s;
  if (r.in_dtype != IN_F32) return MM_ERR_INVALID_INPUT;
  return run_to_buffer(ctx, r, signal_data, signal_length, out);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len, int32_t* n_cu) try {
  Engine* e = Engine::get();
  if (!e) return MM_ERR_METAL_UNAVAILABLE;
  if (name && name_len > 0) {
    std::strncpy(name, e->name().c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (arch && arch_len > 0) {
    std::strncpy(arch, e->arch().c_str(), (size_t)arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  if (n_cu) *n_cu = e->num_cus();
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

}  // extern "C"
