// Audio file ingest: mm_load, mm_get_duration, mm_get_sample_rate.
//
// The reference reads files through AVFoundation (AudioIO.swift:16-147,
// Bridge.swift:1008-1085), which is macOS-only and not part of its sources.
// This file reads RIFF/WAVE files itself: PCM u8 / s16 / s24 / s32 (also in
// WAVE_FORMAT_EXTENSIBLE, samples left-justified in their container) and IEEE
// float32 / float64, any channel count. The frame range, empty results and
// reported rate follow AudioIO.load: start = trunc(offset * file_sr), frames
// to read = trunc(min(available, duration * file_sr)) when duration > 0; no
// frames left -> an empty signal at the output rate. A multichannel file gives
// channel 0 (AudioIO.swift:115-117, 126-128), or the mean of the channels
// when mono is set. A target rate other than the file's is reached with
// mm_resample's converter (Resample.swift) on the GPU, where the raw PCM is
// also converted; a native-rate load converts on the host and needs no GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/metalmom.h"
#include "mm_engine.h"

using namespace mm;

namespace {

struct WavInfo {
  int channels = 0, sr = 0, fmt = 0;
  int64_t data_off = 0, frames = 0;
};

uint32_t le32(const unsigned char* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint16_t le16(const unsigned char* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

// Parses the header of a RIFF/WAVE file; false for anything else.
bool wav_probe(FILE* f, WavInfo* w) {
  if (std::fseek(f, 0, SEEK_END) != 0) return false;
  const int64_t file_size = (int64_t)ftello(f);
  if (std::fseek(f, 0, SEEK_SET) != 0) return false;
  unsigned char h[12];
  if (std::fread(h, 1, 12, f) != 12 || std::memcmp(h, "RIFF", 4) != 0 || std::memcmp(h + 8, "WAVE", 4) != 0)
    return false;
  int64_t pos = 12;
  bool have_fmt = false;
  int bytes = 0, tag = 0, bits = 0;
  while (pos + 8 <= file_size) {
    unsigned char c[8];
    if (fseeko(f, pos, SEEK_SET) != 0 || std::fread(c, 1, 8, f) != 8) return false;
    const uint32_t size = le32(c + 4);
    const int64_t body = pos + 8;
    if (std::memcmp(c, "fmt ", 4) == 0) {
      if (size < 16) return false;
      unsigned char b[40] = {0};
      const size_t n = size < 40 ? size : 40;
      if (std::fread(b, 1, n, f) != n) return false;
      tag = le16(b);
      w->channels = le16(b + 2);
      w->sr = (int)le32(b + 4);
      const int block = le16(b + 12);
      bits = le16(b + 14);
      if (tag == 0xFFFE) {   // WAVE_FORMAT_EXTENSIBLE: the sub-format GUID starts with the tag
        if (n < 26) return false;
        tag = le16(b + 24);
      }
      if (w->channels < 1 || w->sr <= 0 || block % w->channels != 0) return false;
      bytes = block / w->channels;
      have_fmt = true;
    } else if (std::memcmp(c, "data", 4) == 0) {
      if (!have_fmt) return false;
      int64_t n = size;
      if (size == 0 || size == 0xFFFFFFFFu || body + n > file_size) n = file_size - body;   // streamed writers
      if (tag == 1) {
        w->fmt = bytes == 1 ? PCM_U8 : bytes == 2 ? PCM_S16 : bytes == 3 ? PCM_S24 : bytes == 4 ? PCM_S32 : 0;
      } else if (tag == 3) {
        w->fmt = bytes == 4 ? PCM_F32 : bytes == 8 ? PCM_F64 : 0;
      } else {
        w->fmt = 0;
      }
      if (!w->fmt || bits > 8 * bytes) return false;
      w->data_off = body;
      w->frames = n / ((int64_t)bytes * w->channels);
      return true;
    }
    pos = body + size + (size & 1);
  }
  return false;
}

bool wav_open(const char* path, WavInfo* w, FILE** out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  if (!wav_probe(f, w)) {
    std::fclose(f);
    return false;
  }
  *out = f;
  return true;
}

double host_sample(const unsigned char* p, int fmt) {
  switch (fmt) {
    case PCM_U8:
      return ((double)p[0] - 128.0) * (1.0 / 128.0);
    case PCM_S16: {
      int16_t v;
      std::memcpy(&v, p, 2);
      return (double)v * (1.0 / 32768.0);
    }
    case PCM_S24: {
      const int v = (int)p[0] | ((int)p[1] << 8) | ((int)(signed char)p[2] << 16);
      return (double)v * (1.0 / 8388608.0);
    }
    case PCM_S32: {
      int32_t v;
      std::memcpy(&v, p, 4);
      return (double)v * (1.0 / 2147483648.0);
    }
    case PCM_F32: {
      float v;
      std::memcpy(&v, p, 4);
      return (double)v;
    }
    default: {
      double v;
      std::memcpy(&v, p, 8);
      return v;
    }
  }
}

// The host form of k_pcm_to_mono (same operations in the same order).
void host_pcm_to_mono(const unsigned char* raw, int64_t frames, int channels, int fmt, int mix, float* out) {
  const int bytes = pcm_bytes(fmt);
  const int64_t ba = (int64_t)bytes * channels;
  for (int64_t f = 0; f < frames; ++f) {
    const unsigned char* p = raw + f * ba;
    double v = host_sample(p, fmt);
    if (mix && channels > 1) {
      for (int c = 1; c < channels; ++c) v += host_sample(p + c * bytes, fmt);
      v /= (double)channels;
    }
    out[f] = (float)v;
  }
}

void empty_signal(MMBuffer* out) {
  std::memset(out, 0, sizeof(*out));
  out->ndim = 1;   // Bridge.swift:579-595 with a 1-D signal of 0 samples
}

}  // namespace

namespace mm {

int32_t run_pcm_resample(Lane* lane, const void* raw, int64_t frames, int channels, int fmt, int mix,
                         int source_sr, int target_sr, float** out, int64_t* n_out) {
  *out = nullptr;
  const int64_t m = resample_length(frames, source_sr, target_sr);
  *n_out = m;
  if (m <= 0) return OK;
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  hipStream_t st = lane->stream;
  const size_t raw_bytes = (size_t)frames * channels * pcm_bytes(fmt);
  const size_t a = (raw_bytes + 255) & ~(size_t)255, b = ((size_t)frames * 4 + 255) & ~(size_t)255;
  const size_t res_bytes = (size_t)m * 4;
  MM_HIP(lane->h_in.ensure(raw_bytes));
  MM_HIP(lane->d_in.ensure(a + b + res_bytes));   // raw | mono signal | result
  MM_HIP(lane->h_out.ensure(res_bytes));
  std::memcpy(lane->h_in.p, raw, raw_bytes);
  char* d = static_cast<char*>(lane->d_in.p);
  float* mono = reinterpret_cast<float*>(d + a);
  float* res = reinterpret_cast<float*>(d + a + b);
  MM_HIP(hipMemcpyAsync(d, lane->h_in.p, raw_bytes, hipMemcpyHostToDevice, st));
  MM_HIP(launch_pcm_to_mono(d, frames, channels, fmt, mix, mono, e->num_cus(), st));
  const int32_t s = enqueue_resample(mono, frames, source_sr, target_sr, res, st);
  if (s != OK) {
    (void)hipStreamSynchronize(st);
    return s;
  }
  MM_HIP(hipMemcpyAsync(lane->h_out.p, res, res_bytes, hipMemcpyDeviceToHost, st));
  MM_HIP(hipStreamSynchronize(st));
  float* h = static_cast<float*>(std::malloc(res_bytes));
  if (!h) return ERR_ALLOC;
  std::memcpy(h, lane->h_out.p, res_bytes);
  *out = h;
  return OK;
}

}  // namespace mm

extern "C" {

int32_t mm_get_duration(mm_context ctx, const char* path, float* out_duration) try {
  (void)ctx;
  if (!path || !out_duration) return MM_ERR_INVALID_INPUT;
  WavInfo w;
  FILE* f = nullptr;
  if (!wav_open(path, &w, &f)) return MM_ERR_INVALID_INPUT;
  std::fclose(f);
  *out_duration = (float)((double)w.frames / (double)w.sr);   // AudioIO.swift:134-138
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_get_sample_rate(mm_context ctx, const char* path, int32_t* out_sr) try {
  (void)ctx;
  if (!path || !out_sr) return MM_ERR_INVALID_INPUT;
  WavInfo w;
  FILE* f = nullptr;
  if (!wav_open(path, &w, &f)) return MM_ERR_INVALID_INPUT;
  std::fclose(f);
  *out_sr = w.sr;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

static int32_t load_wav(mm_context ctx, const char* path, int32_t sr, int32_t mono, float offset, float duration,
                        MMBuffer* out, int32_t* out_sr);

int32_t mm_load(mm_context ctx, const char* path, int32_t sr, int32_t mono, float offset, float duration,
                MMBuffer* out, int32_t* out_sr) {
  try {
    return load_wav(ctx, path, sr, mono, offset, duration, out, out_sr);
  } catch (const std::bad_alloc&) {   // no exception leaves the C ABI
    return MM_ERR_ALLOC_FAILED;
  }
}

static int32_t load_wav(mm_context ctx, const char* path, int32_t sr, int32_t mono, float offset, float duration,
                        MMBuffer* out, int32_t* out_sr) {
  if (!path || !out) return MM_ERR_INVALID_INPUT;   // Bridge.swift:1058-1060
  if (!(offset >= 0.0f)) return MM_ERR_INVALID_INPUT;
  WavInfo w;
  FILE* f = nullptr;
  if (!wav_open(path, &w, &f)) return MM_ERR_INVALID_INPUT;
  std::unique_ptr<FILE, int (*)(FILE*)> file(f, &std::fclose);   // closed on every path, exceptions included
  const int target = sr > 0 ? sr : w.sr;
  const int64_t start = (int64_t)((double)offset * (double)w.sr);   // AudioIO.swift:31-33
  const int64_t avail = w.frames - start;
  int64_t n = avail;
  if (duration > 0.0f) n = (int64_t)std::min((double)avail, (double)duration * (double)w.sr);
  if (avail <= 0 || n <= 0) {   // AudioIO.swift:34-36
    empty_signal(out);
    if (out_sr) *out_sr = target;
    return MM_OK;
  }
  const int bytes = pcm_bytes(w.fmt);
  const size_t ba = (size_t)bytes * w.channels;
  std::vector<unsigned char> raw((size_t)n * ba);
  const bool ok = fseeko(f, w.data_off + start * (int64_t)ba, SEEK_SET) == 0 &&
                  std::fread(raw.data(), 1, raw.size(), f) == raw.size();
  file.reset();
  if (!ok) return MM_ERR_INVALID_INPUT;
  const int mix = (mono != 0 && w.channels > 1) ? 1 : 0;
  float* data = nullptr;
  int64_t len = n;
  if (target == w.sr) {
    data = static_cast<float*>(std::malloc((size_t)n * sizeof(float)));
    if (!data) return MM_ERR_ALLOC_FAILED;
    host_pcm_to_mono(raw.data(), n, w.channels, w.fmt, mix, data);
  } else {
    Engine* e = Engine::get();
    if (!e) return ERR_NO_GPU;
    Lane* lane = ctx_lane(ctx);
    const bool borrowed = !lane;
    if (!lane) lane = e->acquire_lane();
    if (!lane) return ERR_INTERNAL;
    const int32_t s = run_pcm_resample(lane, raw.data(), n, w.channels, w.fmt, mix, w.sr, target, &data, &len);
    if (borrowed) e->release_lane(lane);
    if (s != OK) return s;
  }
  if (out_sr) *out_sr = target;
  if (len <= 0 || !data) {
    std::free(data);
    empty_signal(out);
    return MM_OK;
  }
  std::memset(out, 0, sizeof(*out));
  out->data = data;
  out->ndim = 1;
  out->shape[0] = len;
  out->count = len;
  return MM_OK;
}

}  // extern "C"
