// Host engine of libmetalmom.so: lazy device singleton, device-resident table
// cache, pooled stream/workspace lanes, batch metadata, fused-kernel and
// generic-kernel execution, and the autotuner's persisted choices.
#include "mm_engine.h"

#include "../../include/metalmom.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include <sys/stat.h>

namespace mm {


// ---------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------

// 1 + (padded - n_fft) / hop, or 0 when the padded signal is shorter than one
// frame (STFT.swift:240-254).
int64_t frame_count(int64_t n, int n_fft, int hop, int center) {
  const int64_t padded = n + (center ? 2 * (int64_t)(n_fft / 2) : 0);
  if (padded < n_fft || hop <= 0) return 0;
  return 1 + (padded - n_fft) / hop;
}

int64_t out_rows(const SpecRequest& r) {
  switch (r.out_kind) {
    case OUT_MEL: return r.n_mels;
    case OUT_MFCC: return std::min(r.n_mfcc, r.n_mels);
    default: return r.n_fft / 2 + 1;
  }
}

int out_elem_floats(const SpecRequest& r) { return r.out_kind == OUT_STFT_COMPLEX ? 2 : 1; }

// ---------------------------------------------------------------------------
// buffers
// ---------------------------------------------------------------------------

hipError_t DevBuf::ensure(size_t bytes) {
  if (bytes <= n && p) return hipSuccess;
  // grow by at least half the old size, so that a stream of growing requests
  // reallocates (and synchronises) a logarithmic number of times
  const size_t old = n;
  release();
  size_t want = std::max<size_t>(std::max<size_t>(bytes, old + old / 2), 256);
  want = (want + (1 << 20) - 1) & ~((size_t)(1 << 20) - 1);
  hipError_t e = hipMalloc(&p, want);
  const size_t exact = (std::max<size_t>(bytes, 1) + (1 << 20) - 1) & ~((size_t)(1 << 20) - 1);
  if (e != hipSuccess && want > exact) {  // no room for the headroom: the exact size
    want = exact;
    e = hipMalloc(&p, exact);
  }
  if (e != hipSuccess) {
    p = nullptr;
    n = 0;
    return e;
  }
  n = want;
  return hipSuccess;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}
hipError_t HostBuf::ensure(size_t bytes) {
  if (bytes <= n && p) return hipSuccess;
  // grow by at least half the old size, so that a stream of growing requests
  // reallocates (and synchronises) a logarithmic number of times
  const size_t old = n;
  release();
  size_t want = std::max<size_t>(std::max<size_t>(bytes, old + old / 2), 4096);
  want = (want + (1 << 20) - 1) & ~((size_t)(1 << 20) - 1);
  hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
  const size_t exact = (std::max<size_t>(bytes, 1) + (1 << 20) - 1) & ~((size_t)(1 << 20) - 1);
  if (e != hipSuccess && want > exact) {  // no room for the headroom: the exact size
    want = exact;
    e = hipHostMalloc(&p, exact, hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    p = nullptr;
    n = 0;
    return e;
  }
  n = want;
  return hipSuccess;
}
void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  n = 0;
}
int32_t Lane::settle() {
  if (busy_set) {
    MM_HIP(hipEventSynchronize(busy));
    busy_set = false;
  }
  return OK;
}

int32_t Lane::enter(hipStream_t st) {
  if (busy_set && st != busy_stream) return settle();
  return OK;
}

int32_t Lane::leave(hipStream_t st) {
  if (!busy) MM_HIP(hipEventCreateWithFlags(&busy, hipEventDisableTiming));
  MM_HIP(hipEventRecord(busy, st));
  busy_stream = st;
  busy_set = true;
  return OK;
}

Lane::~Lane() {
  (void)settle();
  if (busy) (void)hipEventDestroy(busy);
  d_small.release();
  d_in.release();
  d_out.release();
  d_meta.release();
  d_scratch.release();
  d_scratch2.release();
  h_in.release();
  h_out.release();
  h_meta.release();
  h_small.release();
  if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------

struct Tables {
  TableKey key;
  DeviceTables t;
  int cache_chunks[2] = {0, 0};   // filterbank chunks of the LDS cache per chunk width (0: no cache)
  void* mem = nullptr;
  ~Tables() {
    if (mem) (void)hipFree(mem);
  }
};

static TableKey key_of(const SpecRequest& r) {
  TableKey k;
  k.sr = r.sr;
  k.n_fft = r.n_fft;
  k.win = r.win;
  const bool mel = (r.out_kind == OUT_MEL || r.out_kind == OUT_MFCC);
  k.n_mels = mel ? r.n_mels : 0;
  k.n_mfcc = (r.out_kind == OUT_MFCC) ? std::min(r.n_mfcc, r.n_mels) : 0;
  const bool chroma = mel && r.fb_kind == FB_CHROMA;
  k.fmin = (mel && !chroma) ? r.fmin : 0.0f;
  k.fmax = (mel && !chroma) ? r.fmax : 0.0f;
  k.fb_kind = mel ? r.fb_kind : FB_MEL;
  k.tuning = chroma ? r.tuning : 0.0f;
  return k;
}

static int32_t make_tables(const TableKey& key, hipStream_t st, std::unique_ptr<Tables>* out) {
  auto T = std::make_unique<Tables>();
  T->key = key;
  DeviceTables& t = T->t;
  const int M = std::max(key.n_fft / 2, 1);
  const int F = key.n_fft / 2 + 1;
  t.fbs = ((F + 31) / 32) * 32 + 64;
  t.nmb = (key.n_mels + 15) / 16;
  t.nmh = (key.n_mels + kMelBlock - 1) / kMelBlock;
  t.ndt = (key.n_mfcc + 15) / 16;
  const int rows = t.nmb * 16;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_win = take((size_t)key.n_fft * 4);
  const size_t o_tw = take((size_t)M * 8);
  const size_t o_twn = take((size_t)M * 8);
  const size_t o_twp = take((size_t)M * 8);   // per-pass tables total < M entries
  const size_t o_mft = take(key.n_fft == 2048 ? (size_t)kMfTabBytes : 0);
  const size_t o_fb32 = take((size_t)rows * t.fbs * 4);
  const size_t o_fb16 = take((size_t)rows * t.fbs * 2);
  const size_t o_fbh = take((size_t)rows * t.fbs * 2);
  const size_t o_b32 = take((size_t)t.nmh * 8);
  const size_t o_b16 = take((size_t)t.nmh * 8);
  const size_t o_asg = take((size_t)4 * kMaxMelBlocks * 4);
  const size_t o_csrc = take((size_t)2 * kMaxCacheChunks * 4);
  const size_t nd = (size_t)t.ndt * t.nmh * 256;
  const size_t o_d32 = take(nd * 4);
  const size_t o_dbh = take(nd * 2);
  const size_t o_dbl = take(nd * 2);
  const size_t o_dhh = take(nd * 2);
  const size_t o_dhl = take(nd * 2);
  const size_t o_dd = take((size_t)key.n_mfcc * key.n_mels * 4);
  hipError_t e = hipMalloc(&T->mem, off);
  if (e != hipSuccess) return ERR_ALLOC;
  char* b = static_cast<char*>(T->mem);
  t.window = reinterpret_cast<float*>(b + o_win);
  t.tw = reinterpret_cast<float2*>(b + o_tw);
  t.twN = reinterpret_cast<float2*>(b + o_twn);
  t.twp = reinterpret_cast<float2*>(b + o_twp);
  t.mft = (key.n_fft == 2048) ? reinterpret_cast<unsigned char*>(b + o_mft) : nullptr;
  t.fb32 = reinterpret_cast<float*>(b + o_fb32);
  t.fb16 = reinterpret_cast<uint16_t*>(b + o_fb16);
  t.fbh = reinterpret_cast<uint16_t*>(b + o_fbh);
  t.band32 = reinterpret_cast<int2*>(b + o_b32);
  t.band16 = reinterpret_cast<int2*>(b + o_b16);
  t.assign = reinterpret_cast<int*>(b + o_asg);
  t.cache_src = reinterpret_cast<int*>(b + o_csrc);
  t.dct32 = reinterpret_cast<float*>(b + o_d32);
  t.dctb_hi = reinterpret_cast<uint16_t*>(b + o_dbh);
  t.dctb_lo = reinterpret_cast<uint16_t*>(b + o_dbl);
  t.dcth_hi = reinterpret_cast<uint16_t*>(b + o_dhh);
  t.dcth_lo = reinterpret_cast<uint16_t*>(b + o_dhl);
  t.dct_dense = reinterpret_cast<float*>(b + o_dd);
  MM_HIP(gen_tables(key, t, st));
  MM_HIP(hipStreamSynchronize(st));
  if (t.nmh >= 1 && t.nmh <= kMaxMelBlocks) {
    // chunk totals of the two chunk widths (they size the LDS filterbank cache)
    std::vector<int2> bands(2 * (size_t)t.nmh);
    MM_HIP(hipMemcpy(bands.data(), t.band32, (size_t)t.nmh * sizeof(int2), hipMemcpyDeviceToHost));
    MM_HIP(hipMemcpy(bands.data() + t.nmh, t.band16, (size_t)t.nmh * sizeof(int2), hipMemcpyDeviceToHost));
    for (int pc = 0; pc < 2; ++pc) {
      int n = 0;
      for (int b = 0; b < t.nmh; ++b) n += bands[(size_t)pc * t.nmh + b].y;
      T->cache_chunks[pc] = (n <= kMaxCacheChunks) ? n : 0;
    }
  }
  *out = std::move(T);
  return OK;
}

// ---------------------------------------------------------------------------
// engine singleton
// ---------------------------------------------------------------------------

static Engine* g_engine = nullptr;
static int32_t g_status = ERR_NO_GPU;
static std::once_flag g_once;

bool Engine::init() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return false;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return false;
  n_cu_ = p.multiProcessorCount;
  arch_ = p.gcnArchName;
  name_ = p.name;
  // per-block LDS a kernel may request: the larger of the default and the
  // opt-in limit (some runtimes report only 64 KiB as the default)
  int lds = 0, optin = 0;
  (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev);
  const int limit = std::max(lds, optin);
  if (limit > 0) set_lds_limit_bytes((size_t)limit);
  if (const char* v = std::getenv("METALMOM_AMD_TILE_WAVES")) set_tile_waves(std::atoi(v));
  if (const char* v = std::getenv("METALMOM_AMD_VERBOSE"); v && *v == '1') {
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    std::fprintf(stderr, "[metalmom_amd] %s (%s), %d CUs, LDS/block %d (opt-in %d), HIP runtime %d\n",
                 name_.c_str(), arch_.c_str(), n_cu_, lds, optin, rt);
  }
  return true;
}

Engine* Engine::get() {
  std::call_once(g_once, [] {
    Engine* e = new Engine();
    if (e->init()) {
      g_engine = e;
      g_status = OK;
    } else {
      delete e;
      g_status = ERR_NO_GPU;
    }
  });
  return g_engine;
}

int32_t Engine::status() {
  get();
  return g_status;
}

Lane* Engine::acquire_lane() {
  {
    std::lock_guard<std::mutex> g(mu_);
    if (!free_lanes_.empty()) {
      Lane* l = free_lanes_.back();
      free_lanes_.pop_back();
      l->meta_key.clear();   // the previous user may have left another work list in d_meta
      return l;
    }
  }
  Lane* l = new Lane();
  if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) != hipSuccess) {
    delete l;
    return nullptr;
  }
  return l;
}

void Engine::release_lane(Lane* l) {
  if (!l) return;
  (void)l->settle();   // the next borrower may rewrite its buffers at once
  std::lock_guard<std::mutex> g(mu_);
  free_lanes_.push_back(l);
}

int32_t Engine::tables_for(const SpecRequest& r, hipStream_t st, const Tables** out) {
  const TableKey k = key_of(r);
  std::lock_guard<std::mutex> g(mu_);
  for (auto& t : tables_)
    if (t->key == k) {
      *out = t.get();
      return OK;
    }
  std::unique_ptr<Tables> T;
  int32_t s = make_tables(k, st, &T);
  if (s != OK) return s;
  *out = T.get();
  tables_.push_back(std::move(T));
  return OK;
}

const Tables* Engine::tables_resident(const SpecRequest& r) {
  const TableKey k = key_of(r);
  std::lock_guard<std::mutex> g(mu_);
  for (auto& t : tables_)
    if (t->key == k) return t.get();
  return nullptr;
}

int32_t Engine::resample_table(hipStream_t st, const float** out) {
  std::lock_guard<std::mutex> g(mu_);
  if (!rs_table_.p) {
    MM_HIP(rs_table_.ensure((size_t)resample_table_len() * sizeof(float)));
    MM_HIP(launch_tab_resample(static_cast<float*>(rs_table_.p), st));
    MM_HIP(hipStreamSynchronize(st));   // shared by every stream from here on
  }
  *out = static_cast<const float*>(rs_table_.p);
  return OK;
}

int32_t Engine::resample_weights(int source_sr, int target_sr, hipStream_t st, const float** out, float* wsum1) {
  *out = nullptr;
  *wsum1 = 0.0f;
  int P, taps, stride;
  if (!resample_poly_shape(source_sr, target_sr, &P, &taps, &stride)) return OK;
  const float* table = nullptr;
  int32_t s = resample_table(st, &table);
  if (s != OK) return s;
  std::lock_guard<std::mutex> g(mu_);
  for (auto& e : poly_)
    if (e->src == source_sr && e->dst == target_sr && !e->full) {
      *out = static_cast<const float*>(e->w.p);
      *wsum1 = e->wsum1;
      return OK;
    }
  auto e = std::make_unique<PolyEntry>();
  e->src = source_sr;
  e->dst = target_sr;
  e->full = false;
  MM_HIP(e->w.ensure((size_t)(P / 2 + 1) * stride * sizeof(float)));   // rows 0 .. P/2
  MM_HIP(launch_poly_table(source_sr, target_sr, table, static_cast<float*>(e->w.p), st));
  MM_HIP(hipStreamSynchronize(st));   // shared by every stream from here on
  e->wsum1 = 0.0f;
  if (P == 1) {   // the row's sum, in tap order
    std::vector<float> row(taps);
    MM_HIP(hipMemcpy(row.data(), e->w.p, (size_t)taps * sizeof(float), hipMemcpyDeviceToHost));
    for (float v : row) e->wsum1 += v;
  }
  *out = static_cast<const float*>(e->w.p);
  *wsum1 = e->wsum1;
  poly_.push_back(std::move(e));
  return OK;
}

int32_t Engine::resample_phase_weights(int source_sr, int target_sr, hipStream_t st, const float** out) {
  *out = nullptr;
  PhaseShape sh;
  if (!resample_phase_shape(source_sr, target_sr, &sh)) return OK;
  const float* table = nullptr;
  int32_t s = resample_table(st, &table);
  if (s != OK) return s;
  std::lock_guard<std::mutex> g(mu_);
  for (auto& e : poly_)
    if (e->src == source_sr && e->dst == target_sr && e->full) {
      *out = static_cast<const float*>(e->w.p);
      return OK;
    }
  auto e = std::make_unique<PolyEntry>();
  e->src = source_sr;
  e->dst = target_sr;
  e->full = true;
  e->wsum1 = 0.0f;
  MM_HIP(e->w.ensure((size_t)sh.phases * (sh.stride + 1) * sizeof(float)));   // rows, then row sums
  MM_HIP(launch_phase_table(source_sr, target_sr, table, static_cast<float*>(e->w.p), st));
  MM_HIP(hipStreamSynchronize(st));   // shared by every stream from here on
  *out = static_cast<const float*>(e->w.p);
  poly_.push_back(std::move(e));
  return OK;
}

std::string tune_key(const SpecRequest& r, const std::string& arch) {
  std::ostringstream o;
  o << arch << ":k" << r.out_kind << ":n" << r.n_fft << ":h" << r.hop << ":m" << r.n_mels << ":c"
    << ((r.out_kind == OUT_MFCC) ? r.n_mfcc : 0) << ":p" << r.precision << ":i" << r.in_dtype;
  if (r.fb_kind != FB_MEL) o << ":f" << r.fb_kind;
  return o.str();
}

std::string Engine::tuning_path() const {
  const char* p = std::getenv("METALMOM_AMD_TUNE_FILE");
  if (p && *p) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/metalmom_amd_tuning.json";
}

// The file is a flat JSON object {"key": integer, ...}.
void Engine::load_tuning() {
  if (tune_loaded_) return;
  tune_loaded_ = true;
  std::ifstream f(tuning_path());
  if (!f) return;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  size_t i = 0;
  while ((i = s.find('"', i)) != std::string::npos) {
    const size_t j = s.find('"', i + 1);
    if (j == std::string::npos) break;
    const std::string key = s.substr(i + 1, j - i - 1);
    const size_t c = s.find(':', j);
    if (c == std::string::npos) break;
    const int v = std::atoi(s.c_str() + c + 1);
    if (v > 0) tune_.push_back({key, v});
    i = s.find_first_of(",}", c);
    if (i == std::string::npos) break;
  }
}

void Engine::save_tuning() {
  const std::string path = tuning_path();
  const size_t slash = path.find_last_of('/');
  if (slash != std::string::npos && slash > 0) (void)mkdir(path.substr(0, slash).c_str(), 0755);
  std::ofstream f(path);
  if (!f) return;
  f << "{";
  for (size_t i = 0; i < tune_.size(); ++i) f << (i ? ", " : "") << "\"" << tune_[i].first << "\": " << tune_[i].second;
  f << "}\n";
}

// Tuning entries: "<key>:shape" -> the winning TileShape, "<key>:s<shape>"
// -> workgroups per CU for that shape.
static int tune_lookup(const std::vector<std::pair<std::string, int>>& t, const std::string& k) {
  for (auto& e : t)
    if (e.first == k) return e.second;
  return 0;
}

int Engine::tuned_wg_per_cu(const SpecRequest& r, int shape) {
  std::lock_guard<std::mutex> g(mu_);
  load_tuning();
  return tune_lookup(tune_, tune_key(r, arch_) + ":s" + std::to_string(shape));
}

int Engine::tuned_shape(const SpecRequest& r) {
  std::lock_guard<std::mutex> g(mu_);
  load_tuning();
  return tune_lookup(tune_, tune_key(r, arch_) + ":shape");
}

void Engine::record_tuning(const SpecRequest& r, int shape, int wg_per_cu) {
  std::lock_guard<std::mutex> g(mu_);
  load_tuning();
  const std::string base = tune_key(r, arch_);
  auto put = [&](const std::string& k, int v) {
    for (auto& e : tune_)
      if (e.first == k) {
        e.second = v;
        return;
      }
    tune_.push_back({k, v});
  };
  put(base + ":shape", shape);
  put(base + ":s" + std::to_string(shape), wg_per_cu);
  save_tuning();
}

int Engine::grid_for(const SpecRequest& r, int n_tiles, const TileLayout& L) {
  int per_cu = tuned_wg_per_cu(r, L.shape);
  if (per_cu <= 0) {
    // default: as many workgroups as fit by LDS (at most 4 per CU)
    per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, lds_limit_bytes() / std::max<size_t>(L.lds_bytes, 1)));
  }
  const int g = std::max(1, n_cu_ * per_cu);
  return std::min(g, std::max(n_tiles, 1));
}

// ---------------------------------------------------------------------------
// batch metadata
// ---------------------------------------------------------------------------

void build_meta(const SpecRequest& r, const int64_t* lens, int n, int tile_frames, BatchMeta* m) {
  m->n_clips = n;
  m->clip_off.resize(n);
  m->clip_len.resize(n);
  m->out_off.resize(n);
  m->clip_frames.resize(n);
  m->tile_clip.clear();
  m->tile_f0.clear();
  m->tile_pf.clear();
  const int64_t rows = out_rows(r);
  const int ef = out_elem_floats(r);
  int64_t in = 0, out = 0;
  for (int c = 0; c < n; ++c) {
    const int64_t T = frame_count(lens[c], r.n_fft, r.hop, r.center);
    m->clip_off[c] = in;
    m->clip_len[c] = lens[c];
    m->clip_frames[c] = (int32_t)T;
    m->out_off[c] = out;  // in output entries (complex entries for complex STFT)
    in += lens[c];
    out += rows * T;
    const int64_t pad = r.center ? r.n_fft / 2 : 0;
    for (int64_t f = 0; f < T; f += tile_frames) {
      m->tile_clip.push_back(c);
      m->tile_f0.push_back((int32_t)f);
      TilePf p;
      p.base = (long long)(m->clip_off[c] + f * r.hop - pad);
      p.mask = 0u;
      p.pad = 0;
      for (int w = 0; w < tile_frames && w < 32; ++w) {
        const int64_t s0 = (f + w) * r.hop - pad;   // frame start within the clip
        if (f + w < T && s0 >= 0 && s0 + r.n_fft <= lens[c] && (s0 & 1) == 0) p.mask |= 1u << w;
      }
      m->tile_pf.push_back(p);
    }
  }
  m->total_in = in;
  m->total_out = out * ef;
  m->n_tiles = (int)m->tile_clip.size();
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t meta_bytes(const BatchMeta& m) {
  const size_t c = (size_t)std::max(m.n_clips, 1), t = (size_t)std::max(m.n_tiles, 1);
  return align256(c * 8) * 3 + align256(c * 4) + align256(t * 4) * 2 + align256(c * 4) * 2 + align256(t * 4) +
         align256(t * sizeof(TilePf));
}

MetaPtrs meta_ptrs(const BatchMeta& m, void* dev) {
  const size_t c = (size_t)std::max(m.n_clips, 1), t = (size_t)std::max(m.n_tiles, 1);
  char* b = static_cast<char*>(dev);
  MetaPtrs p;
  size_t o = 0;
  p.clip_off = reinterpret_cast<const int64_t*>(b + o); o += align256(c * 8);
  p.clip_len = reinterpret_cast<const int64_t*>(b + o); o += align256(c * 8);
  p.out_off = reinterpret_cast<const int64_t*>(b + o); o += align256(c * 8);
  p.clip_frames = reinterpret_cast<const int32_t*>(b + o); o += align256(c * 4);
  p.tile_clip = reinterpret_cast<const int32_t*>(b + o); o += align256(t * 4);
  p.tile_f0 = reinterpret_cast<const int32_t*>(b + o); o += align256(t * 4);
  p.clip_max_key = reinterpret_cast<unsigned*>(b + o); o += align256(c * 4);
  p.clip_max_key2 = reinterpret_cast<unsigned*>(b + o); o += align256(c * 4);
  p.tile_min = reinterpret_cast<float*>(b + o); o += align256(t * 4);
  p.tile_pf = reinterpret_cast<const TilePf*>(b + o);
  return p;
}

void pack_meta(const BatchMeta& m, void* host) {
  const size_t c = (size_t)std::max(m.n_clips, 1), t = (size_t)std::max(m.n_tiles, 1);
  char* b = static_cast<char*>(host);
  std::memset(b, 0, meta_bytes(m));
  size_t o = 0;
  if (m.n_clips) std::memcpy(b + o, m.clip_off.data(), (size_t)m.n_clips * 8);
  o += align256(c * 8);
  if (m.n_clips) std::memcpy(b + o, m.clip_len.data(), (size_t)m.n_clips * 8);
  o += align256(c * 8);
  if (m.n_clips) std::memcpy(b + o, m.out_off.data(), (size_t)m.n_clips * 8);
  o += align256(c * 8);
  if (m.n_clips) std::memcpy(b + o, m.clip_frames.data(), (size_t)m.n_clips * 4);
  o += align256(c * 4);
  if (m.n_tiles) std::memcpy(b + o, m.tile_clip.data(), (size_t)m.n_tiles * 4);
  o += align256(t * 4);
  if (m.n_tiles) std::memcpy(b + o, m.tile_f0.data(), (size_t)m.n_tiles * 4);
  o += align256(t * 4);
  // the two sets of clip maxima (device-written): -inf, as phase 1 expects them
  for (int k = 0; k < 2; ++k) {
    for (int i = 0; i < m.n_clips; ++i) reinterpret_cast<unsigned*>(b + o)[i] = float_to_key(-INFINITY);
    o += align256(c * 4);
  }
  o += align256(t * 4);   // tile_min (device-written)
  if (m.n_tiles) std::memcpy(b + o, m.tile_pf.data(), (size_t)m.n_tiles * sizeof(TilePf));
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------

// Instrumented builds (-DMM_PHASE_TIMING) write per-wave phase cycles here.
static long long* g_phase_cycles = nullptr;
long long* phase_cycles_buffer() { return g_phase_cycles; }
void set_phase_cycles_buffer(long long* p) { g_phase_cycles = p; }

int launch_grid(Engine* eng, const SpecRequest& r, int n_tiles, const TileLayout& L, int grid_override,
                int* tiles_per_wg) {
  if (n_tiles <= 0) {
    *tiles_per_wg = 0;
    return 0;
  }
  const int grid = grid_override > 0 ? std::min(grid_override, n_tiles) : eng->grid_for(r, n_tiles, L);
  *tiles_per_wg = (n_tiles + grid - 1) / std::max(grid, 1);
  return (n_tiles + *tiles_per_wg - 1) / *tiles_per_wg;
}

static TileArgs make_args(const SpecRequest& r, const Tables* tb, const BatchMeta& m, const MetaPtrs& mp,
                          const void* d_signal, void* d_out, const TileLayout& L, int tiles_per_wg) {
  const DeviceTables& t = tb->t;
  TileArgs a;
  std::memset(&a, 0, sizeof(a));
  a.signal = d_signal;
  a.clip_off = mp.clip_off;
  a.clip_len = mp.clip_len;
  a.clip_frames = mp.clip_frames;
  a.out_off = mp.out_off;
  a.tile_clip = mp.tile_clip;
  a.tile_f0 = mp.tile_f0;
  a.tile_pf = mp.tile_pf;
  a.n_tiles = m.n_tiles;
  a.tiles_per_wg = tiles_per_wg;
  a.n_fft = r.n_fft;
  a.hop = r.hop;
  a.center = r.center;
  a.pad_mode = r.pad_mode;
  a.in_dtype = r.in_dtype;
  a.window = t.window;
  a.tw = t.tw;
  a.twN = t.twN;
  a.twp = t.twp;
  a.mft = t.mft;
  a.fb32 = t.fb32;
  a.fb16 = (r.precision == PREC_F16) ? t.fbh : t.fb16;
  a.band = (r.precision == PREC_F32) ? t.band32 : t.band16;
  // block lists for this chunk width and workgroup shape (8 or 16 waves)
  a.assign = t.assign ? t.assign + (((r.precision == PREC_F32) ? 0 : 2) + (L.nw == 16 ? 1 : 0)) * kMaxMelBlocks
                      : nullptr;
  a.fbs = t.fbs;
  a.nmh = t.nmh;
  a.n_mels = r.n_mels;
  a.dct32 = t.dct32;
  a.dct_hi = (r.precision == PREC_F16) ? t.dcth_hi : t.dctb_hi;
  a.dct_lo = (r.precision == PREC_F16) ? t.dcth_lo : t.dctb_lo;
  a.ndt = t.ndt;
  a.n_mfcc = std::min(r.n_mfcc, r.n_mels);
  a.power = r.power;
  a.post_scale = std::exp2(-r.power);
  a.ref = r.ref;
  a.db_off = (float)(10.0 * std::log10((double)r.ref));
  a.amin = r.amin;
  a.top_db = r.top_db;
  a.use_top_db = (r.out_kind == OUT_MFCC && r.top_db > 0.0f) ? 1 : 0;
  a.phase = 1;
  a.clip_max_key = mp.clip_max_key;
  a.n_clips = m.n_clips;
  a.tile_min = mp.tile_min;
  a.out = d_out;
  a.out_kind = r.out_kind;
  a.lds_row_stride = L.row_stride;
  a.lds_region_bytes = L.region_bytes;
  a.lds_part_off = L.part_off;
  a.lds_stage2_off = L.stage2_off;
  a.lds_part2_off = L.part2_off;
  a.tab_lds_off = L.tab_off;
  a.phase_cycles = phase_cycles_buffer();
  return a;
}

static int32_t enqueue_generic(const SpecRequest& r, const Tables* tb, const BatchMeta& m, const void* d_signal,
                               void* d_out, DevBuf* scratch, DevBuf* scratch2, hipStream_t st) {
  if (r.in_dtype != IN_F32) return ERR_INVALID;
  if (r.n_fft > 16384) return ERR_INVALID;
  const DeviceTables& t = tb->t;
  const int F = r.n_fft / 2 + 1;
  for (int c = 0; c < m.n_clips; ++c) {
    const int T = m.clip_frames[c];
    if (T == 0) continue;
    const float* x = static_cast<const float*>(d_signal) + m.clip_off[c];
    MM_HIP(scratch->ensure((size_t)T * F * 8));
    float2* spec = static_cast<float2*>(scratch->p);
    MM_HIP(launch_generic_fft(x, m.clip_len[c], r.n_fft, r.hop, r.center, r.pad_mode, t.window, t.tw, t.twN, T,
                              spec, st));
    if (r.out_kind <= OUT_STFT_COMPLEX) {
      void* o = static_cast<char*>(d_out) + m.out_off[c] * out_elem_floats(r) * 4;
      MM_HIP(launch_generic_spec_out(spec, F, T, r.out_kind, o, st));
    } else if (r.out_kind == OUT_MEL) {
      float* o = static_cast<float*>(d_out) + m.out_off[c];
      MM_HIP(launch_generic_mel(spec, F, T, t.fb32, t.fbs, r.n_mels, r.power, o, st));
    } else {
      const int nm = std::min(r.n_mfcc, r.n_mels);
      MM_HIP(scratch2->ensure((size_t)r.n_mels * T * 4 + 256));
      float* mel = static_cast<float*>(scratch2->p);
      unsigned* key = reinterpret_cast<unsigned*>(static_cast<char*>(scratch2->p) + (size_t)r.n_mels * T * 4);
      MM_HIP(launch_generic_mel(spec, F, T, t.fb32, t.fbs, r.n_mels, r.power, mel, st));
      MM_HIP(launch_fill_u32(key, 1, float_to_key(-INFINITY), st));
      MM_HIP(launch_db(mel, (int64_t)r.n_mels * T, 10.0f, r.ref, r.amin, mel, key, st));
      if (r.top_db > 0.0f) MM_HIP(launch_db_clamp(mel, (int64_t)r.n_mels * T, key, r.top_db, st));
      float* o = static_cast<float*>(d_out) + m.out_off[c];
      MM_HIP(launch_generic_dct(mel, r.n_mels, T, t.dct_dense, nm, o, st));
    }
  }
  return OK;
}

int32_t prepare_batch(const SpecRequest& r, int shape) {
  TileLayout L;
  if (!tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, std::min(r.n_mfcc, r.n_mels), &L, shape)) return OK;
  TileArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n_fft = r.n_fft;
  a.out_kind = r.out_kind;
  a.in_dtype = r.in_dtype;
  MM_HIP(launch_tile(a, r.precision, L, 0, nullptr));
  return OK;
}

int32_t enqueue_batch(Engine* eng, const SpecRequest& r, const Tables* tb, const BatchMeta& m, const MetaPtrs& mp,
                      const void* d_signal, void* d_out, DevBuf* scratch, DevBuf* scratch2, int grid_override,
                      hipStream_t st, hipEvent_t ev_p1_begin, hipEvent_t ev_p1_end, int shape,
                      unsigned* clear_next) {
  if (m.n_tiles == 0) return OK;
  TileLayout L;
  const bool fused =
      tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, std::min(r.n_mfcc, r.n_mels), &L, shape);
  if (!fused) {
    // a shape the fused kernel does not offer for this configuration
    if (shape != SHAPE_DEFAULT && tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels,
                                              std::min(r.n_mfcc, r.n_mels), &L))
      return ERR_INVALID;
    if (r.precision != PREC_F32) return ERR_INVALID;
    return enqueue_generic(r, tb, m, d_signal, d_out, scratch, scratch2, st);
  }
  int tpw = 0;
  const int g = launch_grid(eng, r, m.n_tiles, L, grid_override, &tpw);
  TileArgs a = make_args(r, tb, m, mp, d_signal, d_out, L, tpw);
  // LDS copies of the filterbank band and the DCT operands, when they fit
  // beside the kernel's own layout
  TileLayout LL = L;
  if (r.out_kind == OUT_MEL || r.out_kind == OUT_MFCC) {
    const int pc = (r.precision == PREC_F32) ? 0 : 1;
    const int nch = tb->cache_chunks[pc];
    const size_t off = (L.lds_bytes + 255) / 256 * 256;
    const size_t dcb = (r.out_kind == OUT_MFCC) ? (size_t)tb->t.ndt * tb->t.nmh * 1024 : 0;
    // (within the LDS share of the workgroups per CU the shape is sized for):
    // the DCT operands (read by every wave in every tile) when they fit, and
    // beside them the filterbank chunks in block order, all of them or as
    // many whole blocks' worth as fit (the kernel reads a block from LDS when
    // all of its chunks are cached)
    const size_t lim = lds_limit_bytes() / std::max(L.occ, 1);
    if (L.stage2_off > 0) {
      // pipelined layout: the filterbank chunks first (every wave reads its
      // band in every tile), the DCT operands only if they still fit
      const int nc = (int)std::min<size_t>((size_t)nch, off < lim ? (lim - off) / kFbChunkBytes : 0);
      if (nc > 0) {
        a.fb_lds_off = (int)off;
        a.cache_src = tb->t.cache_src + pc * kMaxCacheChunks;
        a.cache_chunks = nc;
      }
      const size_t fbc = (size_t)nc * kFbChunkBytes;
      const bool dfit = dcb > 0 && off + fbc + dcb <= lim;
      a.dct_lds_off = dfit ? (int)(off + fbc) : 0;
      LL.lds_bytes = off + fbc + (dfit ? dcb : 0);
    } else if (off < lim && (nch > 0 || dcb > 0)) {
      // DCT operands too large for the room left (n_fft 4096 with 256 mels):
      // the filterbank chunks alone, the DCT operands from L2
      const bool dfit = off + dcb <= lim;
      const size_t dct_bytes = dfit ? dcb : 0;
      const int nc = (int)std::min<size_t>((size_t)nch, (lim - off - dct_bytes) / kFbChunkBytes);
      if (nc > 0) {
        a.fb_lds_off = (int)off;
        a.cache_src = tb->t.cache_src + pc * kMaxCacheChunks;
        a.cache_chunks = nc;
      }
      const size_t fbc = (size_t)nc * kFbChunkBytes;
      a.dct_lds_off = (dcb && dfit) ? (int)(off + fbc) : 0;
      LL.lds_bytes = off + fbc + dct_bytes;
    }
  }
  // clear_next: mp.clip_max_key is at -inf already (a plan's alternating
  // sets: the previous step's phase 1 cleared it) and this phase 1 clears
  // the other set for the next step; otherwise a fill launch clears it
  if (a.use_top_db) {
    if (clear_next) a.clip_max_clear = clear_next;
    else MM_HIP(launch_fill_u32(mp.clip_max_key, m.n_clips, float_to_key(-INFINITY), st));
  }
  if (ev_p1_begin) MM_HIP(hipEventRecord(ev_p1_begin, st));
  MM_HIP(launch_tile(a, r.precision, LL, g, st));
  if (ev_p1_end) MM_HIP(hipEventRecord(ev_p1_end, st));
  if (a.use_top_db) {
    a.phase = 2;
    MM_HIP(launch_tile(a, r.precision, LL, g, st));
  }
  return OK;
}

int32_t run_single(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, float** out, int64_t* rows,
                   int64_t* cols) {
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  hipStream_t st = lane->stream;
  const Tables* tb = nullptr;
  int32_t s = eng->tables_for(r, st, &tb);
  if (s != OK) return s;
  TileLayout L;
  const bool fused = tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, std::min(r.n_mfcc, r.n_mels), &L);
  BatchMeta m;
  build_meta(r, &len, 1, fused ? L.tile_frames : kTileFrames, &m);
  *rows = out_rows(r);
  *cols = m.clip_frames[0];
  *out = nullptr;
  const size_t out_bytes = (size_t)m.total_out * 4;
  if (out_bytes == 0) return OK;
  const size_t in_bytes = (size_t)len * 4;
  const size_t mb = meta_bytes(m);
  MM_HIP(lane->h_in.ensure(in_bytes + mb));
  MM_HIP(lane->d_in.ensure(in_bytes));
  MM_HIP(lane->d_meta.ensure(mb));
  lane->meta_key.clear();   // d_meta now holds this call's work list
  MM_HIP(lane->d_out.ensure(out_bytes));
  MM_HIP(lane->h_out.ensure(out_bytes));
  std::memcpy(lane->h_in.p, signal, in_bytes);
  pack_meta(m, static_cast<char*>(lane->h_in.p) + in_bytes);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, in_bytes, hipMemcpyHostToDevice, st));
  MM_HIP(hipMemcpyAsync(lane->d_meta.p, static_cast<char*>(lane->h_in.p) + in_bytes, mb, hipMemcpyHostToDevice, st));
  const MetaPtrs mp = meta_ptrs(m, lane->d_meta.p);
  s = enqueue_batch(eng, r, tb, m, mp, lane->d_in.p, lane->d_out.p, &lane->d_scratch, &lane->d_scratch2, 0, st);
  if (s != OK) {
    (void)hipStreamSynchronize(st);
    return s;
  }
  MM_HIP(hipMemcpyAsync(lane->h_out.p, lane->d_out.p, out_bytes, hipMemcpyDeviceToHost, st));
  MM_HIP(hipStreamSynchronize(st));
  float* res = static_cast<float*>(std::malloc(out_bytes));
  if (!res) return ERR_ALLOC;
  std::memcpy(res, lane->h_out.p, out_bytes);
  *out = res;
  return OK;
}

// the clip's work list in the lane's d_meta, uploaded once per (length, request)
static int32_t cached_meta(Lane* lane, const SpecRequest& r, int64_t len, int tile_frames, const BatchMeta& m,
                           hipStream_t st) {
  const std::string key = std::to_string(len) + ":" + tune_key(r, "") + ":" + std::to_string(r.center) + ":" +
                          std::to_string(r.win) + ":" + std::to_string(r.sr) + ":" + std::to_string(r.fmin) + ":" +
                          std::to_string(r.fmax) + ":" + std::to_string(r.tuning) + ":" + std::to_string(tile_frames);
  if (lane->meta_key == key) return OK;
  const size_t mb = meta_bytes(m);
  MM_HIP(lane->h_meta.ensure(mb));
  MM_HIP(lane->d_meta.ensure(mb));
  pack_meta(m, lane->h_meta.p);
  MM_HIP(hipMemcpyAsync(lane->d_meta.p, lane->h_meta.p, mb, hipMemcpyHostToDevice, st));
  MM_HIP(hipStreamSynchronize(st));
  lane->meta_key = key;
  return OK;
}

int32_t enqueue_descriptor(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int kind, float param,
                           float* d_res, hipStream_t st) {
  const int T = (int)frame_count(len, r.n_fft, r.hop, r.center);
  if (T == 0) return OK;
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  const Tables* tb = nullptr;
  int32_t s = eng->tables_for(r, st, &tb);
  if (s != OK) return s;
  TileLayout L;
  const bool fused = tile_layout(r.n_fft, r.out_kind, r.precision, 0, 0, &L);
  const int nw = desc_fused_waves(r.n_fft);
  if (fused && nw > 0 && r.in_dtype == IN_F32) {
    // FFT and reduction in one kernel: one frame per wave, tiles of nw frames
    BatchMeta m;
    build_meta(r, &len, 1, nw, &m);
    s = cached_meta(lane, r, len, nw, m, st);
    if (s != OK) return s;
    const MetaPtrs mp = meta_ptrs(m, lane->d_meta.p);
    const int grid = std::max(1, std::min(m.n_tiles, eng->num_cus()));
    const TileArgs a = make_args(r, tb, m, mp, d_x, d_res, L, (m.n_tiles + grid - 1) / grid);
    const int g = (m.n_tiles + a.tiles_per_wg - 1) / a.tiles_per_wg;
    MM_HIP(launch_desc_fused(a, kind, (float)r.sr, (float)r.n_fft, param, d_res, g, st));
    return OK;
  }
  const int tf = fused ? L.tile_frames : kTileFrames;
  BatchMeta m;
  build_meta(r, &len, 1, tf, &m);
  MM_HIP(lane->d_out.ensure((size_t)m.total_out * 4));   // magnitude spectrogram [F, T]
  s = cached_meta(lane, r, len, tf, m, st);
  if (s != OK) return s;
  const MetaPtrs mp = meta_ptrs(m, lane->d_meta.p);
  s = enqueue_batch(eng, r, tb, m, mp, d_x, lane->d_out.p, &lane->d_scratch, &lane->d_scratch2, 0, st);
  if (s != OK) return s;
  MM_HIP(launch_descriptor(static_cast<const float*>(lane->d_out.p), r.n_fft / 2 + 1, T, kind, (float)r.sr,
                           (float)r.n_fft, param, d_res, st));
  return OK;
}

int32_t run_descriptor(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int kind, float param,
                       float** out, int64_t* frames) {
  *out = nullptr;
  const int T = (int)frame_count(len, r.n_fft, r.hop, r.center);
  *frames = T;
  if (T == 0) return OK;
  hipStream_t st = lane->stream;
  const size_t in_bytes = (size_t)len * 4, res_off = (in_bytes + 255) & ~(size_t)255;
  const size_t res_bytes = (size_t)T * 4;
  MM_HIP(lane->h_in.ensure(in_bytes));
  MM_HIP(lane->d_in.ensure(res_off + res_bytes));   // signal | result
  MM_HIP(lane->h_out.ensure(res_bytes));
  std::memcpy(lane->h_in.p, signal, in_bytes);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, in_bytes, hipMemcpyHostToDevice, st));
  float* res = reinterpret_cast<float*>(static_cast<char*>(lane->d_in.p) + res_off);
  const int32_t s = enqueue_descriptor(lane, r, static_cast<const float*>(lane->d_in.p), len, kind, param, res, st);
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

int64_t resample_length(int64_t n, int source_sr, int target_sr) {
  if (source_sr == target_sr) return n;
  if (n <= 0) return 0;
  return (int64_t)std::ceil((double)n * ((double)target_sr / (double)source_sr));   // Resample.swift:21-22
}

int32_t enqueue_resample(const float* d_x, int64_t n, int source_sr, int target_sr, float* d_y, hipStream_t st) {
  if (n <= 0) return OK;
  if (source_sr == target_sr) {   // Resample.swift:13: the signal unchanged
    MM_HIP(hipMemcpyAsync(d_y, d_x, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st));
    return OK;
  }
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  const int64_t m = resample_length(n, source_sr, target_sr);
  const float* w = nullptr;
  int32_t s = e->resample_phase_weights(source_sr, target_sr, st, &w);
  if (s != OK) return s;
  if (w) {   // two or more phases: one phase per wave, weight rows by scalar loads
    MM_HIP(launch_resample_phase(d_x, n, d_y, m, source_sr, target_sr, w, e->num_cus(), st));
    return OK;
  }
  float wsum1 = 0.0f;
  s = e->resample_weights(source_sr, target_sr, st, &w, &wsum1);
  if (s != OK) return s;
  if (w) {   // polyphase: weight rows in LDS (one phase: the row by scalar loads)
    if (source_sr == 2 * target_sr) {   // 256 taps: always fits the two-outputs-per-lane kernel
      MM_HIP(launch_resample_dec2(d_x, n, d_y, m, source_sr, target_sr, w, wsum1, e->num_cus(), st));
      return OK;
    }
    MM_HIP(launch_resample_poly(d_x, n, d_y, m, source_sr, target_sr, w, wsum1, e->num_cus(), st));
    return OK;
  }
  const float* table = nullptr;
  s = e->resample_table(st, &table);
  if (s != OK) return s;
  MM_HIP(launch_resample(d_x, n, d_y, m, source_sr, target_sr, table, e->num_cus(), st));
  return OK;
}

int32_t run_resample(Lane* lane, const float* x, int64_t n, int source_sr, int target_sr, float** out,
                     int64_t* n_out) {
  *out = nullptr;
  const int64_t m = resample_length(n, source_sr, target_sr);
  *n_out = m;
  if (m <= 0) return OK;
  hipStream_t st = lane->stream;
  const size_t in_bytes = (size_t)n * 4, res_off = (in_bytes + 255) & ~(size_t)255;
  const size_t res_bytes = (size_t)m * 4;
  MM_HIP(lane->h_in.ensure(in_bytes));
  MM_HIP(lane->d_in.ensure(res_off + res_bytes));   // signal | result
  MM_HIP(lane->h_out.ensure(res_bytes));
  std::memcpy(lane->h_in.p, x, in_bytes);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, in_bytes, hipMemcpyHostToDevice, st));
  float* res = reinterpret_cast<float*>(static_cast<char*>(lane->d_in.p) + res_off);
  const int32_t s = enqueue_resample(static_cast<const float*>(lane->d_in.p), n, source_sr, target_sr, res, st);
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

bool contrast_bands(int sr, int n_fft, int n_freqs, int n_bands, float fmin, ContrastBands* cb) {
  if (n_bands < 1 || n_bands + 1 > kMaxContrastBands) return false;
  // float32 as SpectralDescriptors.swift:145-158 forms them; roundf rounds half away from zero
  const float res = (float)sr / (float)n_fft;
  cb->nb = n_bands + 1;
  for (int b = 0; b <= n_bands + 1; ++b) {
    const float hz = (b == 0) ? 0.0f : (b <= n_bands) ? fmin * powf(2.0f, (float)(b - 1)) : (float)sr / 2.0f;
    cb->edge[b] = std::min((int)roundf(hz / res), n_freqs);
  }
  for (int b = 0; b <= n_bands; ++b) {
    const int n = cb->edge[b + 1] - cb->edge[b];
    cb->q[b] = n > 0 ? std::max(1, (int)roundf(0.02f * (float)n)) : 0;   // quantile 0.02 (:180-181)
  }
  return true;
}

int32_t enqueue_contrast(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int n_bands, float fmin,
                         float* d_res, hipStream_t st) {
  const int T = (int)frame_count(len, r.n_fft, r.hop, r.center);
  if (T == 0) return OK;
  ContrastBands cb;
  if (!contrast_bands(r.sr, r.n_fft, r.n_fft / 2 + 1, n_bands, fmin, &cb)) return ERR_INVALID;
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  const Tables* tb = nullptr;
  int32_t s = eng->tables_for(r, st, &tb);
  if (s != OK) return s;
  TileLayout L;
  const bool fused = tile_layout(r.n_fft, r.out_kind, r.precision, 0, 0, &L);
  const int tf = fused ? L.tile_frames : kTileFrames;
  BatchMeta m;
  build_meta(r, &len, 1, tf, &m);
  MM_HIP(lane->d_out.ensure((size_t)m.total_out * 4));   // magnitude spectrogram [F, T]
  s = cached_meta(lane, r, len, tf, m, st);
  if (s != OK) return s;
  // band table in d_small after the onset key slot, uploaded once per configuration
  constexpr size_t kBandsOff = 256;
  const void* before = lane->d_small.p;
  MM_HIP(lane->d_small.ensure(kBandsOff + sizeof(ContrastBands)));
  if (lane->d_small.p != before) lane->bands_key.clear();
  unsigned fbits;
  std::memcpy(&fbits, &fmin, 4);
  const std::string key = std::to_string(r.sr) + ":" + std::to_string(r.n_fft) + ":" + std::to_string(n_bands) +
                          ":" + std::to_string(fbits);
  auto* d_cb = reinterpret_cast<const ContrastBands*>(static_cast<char*>(lane->d_small.p) + kBandsOff);
  if (lane->bands_key != key) {
    MM_HIP(lane->h_small.ensure(sizeof(ContrastBands)));
    std::memcpy(lane->h_small.p, &cb, sizeof(ContrastBands));
    MM_HIP(hipMemcpyAsync(const_cast<ContrastBands*>(d_cb), lane->h_small.p, sizeof(ContrastBands),
                          hipMemcpyHostToDevice, st));
    MM_HIP(hipStreamSynchronize(st));
    lane->bands_key = key;
  }
  const MetaPtrs mp = meta_ptrs(m, lane->d_meta.p);
  s = enqueue_batch(eng, r, tb, m, mp, d_x, lane->d_out.p, &lane->d_scratch, &lane->d_scratch2, 0, st);
  if (s != OK) return s;
  MM_HIP(launch_contrast(static_cast<const float*>(lane->d_out.p), T, cb.nb, d_cb, d_res, st));
  return OK;
}

int32_t run_contrast(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int n_bands, float fmin,
                     float** out, int64_t* frames) {
  *out = nullptr;
  const int T = (int)frame_count(len, r.n_fft, r.hop, r.center);
  *frames = T;
  if (T == 0) return OK;
  hipStream_t st = lane->stream;
  const size_t in_bytes = (size_t)len * 4, res_off = (in_bytes + 255) & ~(size_t)255;
  const size_t res_bytes = (size_t)(n_bands + 1) * T * 4;
  MM_HIP(lane->h_in.ensure(in_bytes));
  MM_HIP(lane->d_in.ensure(res_off + res_bytes));   // signal | result
  MM_HIP(lane->h_out.ensure(res_bytes));
  std::memcpy(lane->h_in.p, signal, in_bytes);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, in_bytes, hipMemcpyHostToDevice, st));
  float* res = reinterpret_cast<float*>(static_cast<char*>(lane->d_in.p) + res_off);
  const int32_t s = enqueue_contrast(lane, r, static_cast<const float*>(lane->d_in.p), len, n_bands, fmin, res, st);
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

OnsetShape onset_shape(const SpecRequest& r, int64_t len, int aggregate, int lag) {
  OnsetShape o;
  o.T = (int)frame_count(len, r.n_fft, r.hop, r.center);
  o.rows = aggregate ? 1 : r.n_mels;
  if (o.T <= lag) {   // OnsetDetection.swift:63-71: one zero column
    o.cols = 1;
    return o;
  }
  o.pad = lag + (r.center ? r.n_fft / (2 * r.hop) : 0);
  o.cols = r.center ? std::min(o.T - lag + o.pad, o.T) : o.T - lag + o.pad;
  return o;
}

int32_t enqueue_onset(Lane* lane, const SpecRequest& r, const float* d_x, int64_t len, int aggregate, int lag,
                      float* d_res, hipStream_t st) {
  const OnsetShape o = onset_shape(r, len, aggregate, lag);
  if (o.T <= lag) {
    MM_HIP(launch_fill_u32(reinterpret_cast<unsigned*>(d_res), o.rows, 0u, st));
    return OK;
  }
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  const Tables* tb = nullptr;
  int32_t s = eng->tables_for(r, st, &tb);
  if (s != OK) return s;
  TileLayout L;
  const bool fused = tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, 0, &L);
  BatchMeta m;
  build_meta(r, &len, 1, fused ? L.tile_frames : kTileFrames, &m);
  // every buffer is sized before the first launch
  MM_HIP(lane->d_out.ensure((size_t)m.total_out * 4));   // mel, then its dB in place
  MM_HIP(lane->d_small.ensure(256));                      // maximum key
  s = cached_meta(lane, r, len, fused ? L.tile_frames : kTileFrames, m, st);
  if (s != OK) return s;
  const MetaPtrs mp = meta_ptrs(m, lane->d_meta.p);
  s = enqueue_batch(eng, r, tb, m, mp, d_x, lane->d_out.p, &lane->d_scratch, &lane->d_scratch2, 0, st);
  if (s != OK) return s;
  float* mel = static_cast<float*>(lane->d_out.p);
  unsigned* mkey = static_cast<unsigned*>(lane->d_small.p);
  MM_HIP(launch_fill_u32(mkey, 1, float_to_key(-INFINITY), st));
  MM_HIP(launch_db(mel, (int64_t)r.n_mels * o.T, 10.0f, 1.0f, 1e-10f, mel, mkey, st));   // Scaling.swift:73-114
  MM_HIP(launch_onset(mel, r.n_mels, o.T, mkey, 80.0f, lag, o.pad, o.cols, aggregate, d_res, st));
  return OK;
}

int32_t run_onset(Lane* lane, const SpecRequest& r, const float* signal, int64_t len, int aggregate, int lag,
                  float** out, int64_t* rows, int64_t* cols) {
  *out = nullptr;
  const OnsetShape o = onset_shape(r, len, aggregate, lag);
  *rows = o.rows;
  *cols = o.cols;
  const size_t res_bytes = (size_t)o.rows * o.cols * 4;
  if (o.T <= lag) {
    float* z = static_cast<float*>(std::calloc((size_t)o.rows, sizeof(float)));
    if (!z) return ERR_ALLOC;
    *out = z;
    return OK;
  }
  hipStream_t st = lane->stream;
  const size_t in_bytes = (size_t)len * 4, res_off = (in_bytes + 255) & ~(size_t)255;
  MM_HIP(lane->h_in.ensure(in_bytes));
  MM_HIP(lane->d_in.ensure(res_off + res_bytes));   // signal | result
  MM_HIP(lane->h_out.ensure(res_bytes));
  std::memcpy(lane->h_in.p, signal, in_bytes);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, in_bytes, hipMemcpyHostToDevice, st));
  float* res = reinterpret_cast<float*>(static_cast<char*>(lane->d_in.p) + res_off);
  const int32_t s = enqueue_onset(lane, r, static_cast<const float*>(lane->d_in.p), len, aggregate, lag, res, st);
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

int32_t run_db(Lane* lane, const float* data, int64_t n, float mult, float ref, float amin, float top_db,
               float** out) {
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  hipStream_t st = lane->stream;
  const size_t bytes = (size_t)n * 4;
  const size_t key_off = (bytes + 255) & ~(size_t)255;
  MM_HIP(lane->h_in.ensure(bytes));
  MM_HIP(lane->d_in.ensure(key_off + 256));
  MM_HIP(lane->d_out.ensure(bytes));
  MM_HIP(lane->h_out.ensure(bytes));
  std::memcpy(lane->h_in.p, data, bytes);
  unsigned* key = reinterpret_cast<unsigned*>(static_cast<char*>(lane->d_in.p) + key_off);
  MM_HIP(hipMemcpyAsync(lane->d_in.p, lane->h_in.p, bytes, hipMemcpyHostToDevice, st));
  MM_HIP(launch_fill_u32(key, 1, float_to_key(-INFINITY), st));
  float* d = static_cast<float*>(lane->d_out.p);
  MM_HIP(launch_db(static_cast<const float*>(lane->d_in.p), n, mult, ref, amin, d, key, st));
  if (top_db > 0.0f) MM_HIP(launch_db_clamp(d, n, key, top_db, st));
  MM_HIP(hipMemcpyAsync(lane->h_out.p, d, bytes, hipMemcpyDeviceToHost, st));
  MM_HIP(hipStreamSynchronize(st));
  float* res = static_cast<float*>(std::malloc(bytes));
  if (!res) return ERR_ALLOC;
  std::memcpy(res, lane->h_out.p, bytes);
  *out = res;
  return OK;
}

// ---------------------------------------------------------------------------
// additive API requests
// ---------------------------------------------------------------------------

int32_t request_from_params(const MMSpecParams* p, SpecRequest* r) {
  if (!p) return ERR_INVALID;
  const int n_fft = p->n_fft;
  if (n_fft <= 0 || (n_fft & (n_fft - 1))) return ERR_INVALID;
  if (p->kind < MM_KIND_STFT_MAG || p->kind > MM_KIND_CHROMA) return ERR_INVALID;
  const bool chroma = p->kind == MM_KIND_CHROMA;
  r->out_kind = chroma ? OUT_MEL : p->kind;   // chroma: the banded contraction with the chroma table
  r->fb_kind = chroma ? FB_CHROMA : FB_MEL;
  r->tuning = 0.0f;
  r->sr = p->sample_rate;
  r->n_fft = n_fft;
  r->hop = p->hop_length > 0 ? p->hop_length : n_fft / 4;
  r->win = p->win_length > 0 ? p->win_length : n_fft;
  r->center = p->center != 0;
  r->pad_mode = p->pad_mode;
  r->power = p->power;
  r->n_mels = p->n_mels;
  r->fmin = p->f_min;
  r->fmax = p->f_max > 0.0f ? p->f_max : (float)p->sample_rate / 2.0f;
  r->n_mfcc = p->n_mfcc;
  r->top_db = p->top_db;
  r->precision = p->precision;
  r->in_dtype = p->input_dtype;
  if (r->hop <= 0) return ERR_INVALID;
  if (r->pad_mode != PAD_CONSTANT && r->pad_mode != PAD_REFLECT) return ERR_INVALID;
  if (r->precision < PREC_F32 || r->precision > PREC_F16) return ERR_INVALID;
  if (r->in_dtype != IN_F32 && r->in_dtype != IN_F16) return ERR_INVALID;
  if (r->out_kind >= OUT_MEL && r->n_mels < 0) return ERR_INVALID;
  if (chroma && (r->n_mels < 1 || r->n_mels > 256 || r->sr <= 0)) return ERR_INVALID;
  if (r->out_kind == OUT_MFCC && r->n_mfcc < 0) return ERR_INVALID;
  if (r->out_kind <= OUT_STFT_COMPLEX && r->precision != PREC_F32) return ERR_INVALID;
  if (r->precision != PREC_F32 || r->in_dtype != IN_F32) {
    // 16-bit contractions and FP16 input exist only in the fused kernel
    TileLayout L;
    if (!tile_layout(r->n_fft, r->out_kind, r->precision, r->n_mels, std::min(r->n_mfcc, r->n_mels), &L))
      return ERR_INVALID;
  }
  return OK;
}

// ---------------------------------------------------------------------------
// inverse STFT
// ---------------------------------------------------------------------------

int32_t istft_shape(int n_freqs, int n_frames, int hop, int center, int64_t length, IstftShape* sh) {
  if (n_freqs < 1 || n_frames < 1 || hop < 0) return ERR_INVALID;
  IstftShape s;
  s.n_fft = (n_freqs - 1) * 2;
  s.hop = hop;
  s.T = n_frames;
  s.total = (int64_t)s.n_fft + (int64_t)(n_frames - 1) * hop;
  s.pow2 = s.n_fft > 0 && (s.n_fft & (s.n_fft - 1)) == 0;
  if (!s.pow2) {   // STFT.swift:720-724: the zero buffer itself
    s.start = 0;
    s.out_len = s.total;
  } else {
    if (s.n_fft > 16384) return ERR_INVALID;
    s.start = center ? s.n_fft / 2 : 0;
    if (length > 0) s.out_len = length;
    else if (center) s.out_len = (s.total - s.n_fft / 2 > s.n_fft / 2) ? s.total - s.n_fft : 0;
    else s.out_len = s.total;
  }
  if (s.total >= ((int64_t)1 << 31) || s.out_len >= ((int64_t)1 << 31)) return ERR_INVALID;
  *sh = s;
  return OK;
}

int32_t enqueue_istft(const IstftShape& sh, int sr, int win, const float* d_X, int ld, float* d_out,
                      DevBuf* scratch, hipStream_t st) {
  Engine* eng = Engine::get();
  if (!eng) return ERR_NO_GPU;
  if (sh.out_len == 0) return OK;
  if (!sh.pow2) {
    MM_HIP(launch_fill_u32(reinterpret_cast<unsigned*>(d_out), sh.out_len, 0u, st));
    return OK;
  }
  SpecRequest r;
  r.out_kind = OUT_STFT_COMPLEX;
  r.sr = sr;
  r.n_fft = sh.n_fft;
  r.hop = sh.hop > 0 ? sh.hop : 1;
  r.win = win;
  const Tables* tb = nullptr;
  int32_t s = eng->tables_for(r, st, &tb);
  if (s != OK) return s;
  IstftArgs a{};
  a.X = reinterpret_cast<const float2*>(d_X);
  a.out = d_out;
  a.window = tb->t.window;
  a.tw = tb->t.tw;
  a.twN = tb->t.twN;
  a.twp = tb->t.twp;
  a.n_fft = sh.n_fft;
  a.hop = sh.hop;
  a.T = sh.T;
  a.ld = ld;
  a.total = sh.total;
  a.start = sh.start;
  a.out_len = sh.out_len;
  IstftPlan pl;
  if (istft_fused_plan(sh.n_fft, sh.hop, &pl)) {
    const int64_t span = (int64_t)pl.nw * sh.hop;   // output samples per tile
    const int64_t t0 = sh.start / span;
    const int64_t t1 = (sh.start + sh.out_len + span - 1) / span;
    const int n_tiles = (int)(t1 - t0);
    const int per_cu =
        (int)std::max<size_t>(1, std::min<size_t>(4, lds_limit_bytes() / std::max<size_t>(pl.lds_bytes, 1)));
    int grid = std::min(n_tiles, std::max(1, eng->num_cus() * per_cu));
    const int per_wg = (n_tiles + grid - 1) / grid;
    grid = (n_tiles + per_wg - 1) / per_wg;
    a.halo = pl.halo;
    a.carry = pl.carry;
    a.carry_stride = pl.carry_stride;
    a.tile0 = (int)t0;
    a.n_tiles = n_tiles;
    a.tiles_per_wg = per_wg;
    a.region_bytes = pl.region_bytes;
    MM_HIP(launch_istft_tile(a, pl.nw, grid, pl.lds_bytes, st));
    return OK;
  }
  MM_HIP(scratch->ensure((size_t)sh.T * sh.n_fft * sizeof(float)));
  MM_HIP(launch_istft_generic(a, static_cast<float*>(scratch->p), st));
  return OK;
}

int32_t run_istft(Lane* lane, const IstftShape& sh, int sr, int win, const float* X, float** out) {
  *out = nullptr;
  if (sh.out_len == 0) return OK;
  hipStream_t st = lane->stream;
  // device rows padded to a multiple of 16 frames: the fused kernel then
  // reads whole 128-byte lines
  const int F = sh.n_fft / 2 + 1;
  const int ld = (sh.T + 15) / 16 * 16;
  const size_t row = (size_t)sh.T * 8, drow = (size_t)ld * 8;
  const size_t in_bytes = row * F;
  const size_t out_bytes = (size_t)sh.out_len * sizeof(float);
  MM_HIP(lane->h_in.ensure(in_bytes));
  MM_HIP(lane->d_in.ensure(drow * F));
  MM_HIP(lane->d_out.ensure(out_bytes));
  MM_HIP(lane->h_out.ensure(out_bytes));
  std::memcpy(lane->h_in.p, X, in_bytes);
  MM_HIP(hipMemcpy2DAsync(lane->d_in.p, drow, lane->h_in.p, row, row, F, hipMemcpyHostToDevice, st));
  const int32_t s = enqueue_istft(sh, sr, win, static_cast<const float*>(lane->d_in.p), ld,
                                  static_cast<float*>(lane->d_out.p), &lane->d_scratch, st);
  if (s != OK) {
    (void)hipStreamSynchronize(st);
    return s;
  }
  MM_HIP(hipMemcpyAsync(lane->h_out.p, lane->d_out.p, out_bytes, hipMemcpyDeviceToHost, st));
  MM_HIP(hipStreamSynchronize(st));
  float* res = static_cast<float*>(std::malloc(out_bytes));
  if (!res) return ERR_ALLOC;
  std::memcpy(res, lane->h_out.p, out_bytes);
  *out = res;
  return OK;
}

}  // namespace mm
