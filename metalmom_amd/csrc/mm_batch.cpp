// Batched execution API (additive exports of include/metalmom.h): device-
// resident batch plans with one persistent launch per step, hipGraph replay,
// the grid autotuner and event timing, plus the host-buffer batch call.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/metalmom.h"
#include "mm_engine.h"

using namespace mm;

namespace {

// Captured steps of one (input, output, stream), one per set of top_db clip
// maxima (steps alternate between the two).
struct GraphEntry {
  const void* in;
  float* out;
  hipStream_t st;
  hipGraph_t graph[2];
  hipGraphExec_t exec[2];
};

struct Batch {
  SpecRequest r;
  BatchMeta m;
  const Tables* tb = nullptr;
  Lane* lane = nullptr;  // own stream, metadata and scratch
  MetaPtrs mp{};
  int wg_per_cu = 0;
  int shape = SHAPE_DEFAULT;   // workgroup shape the tile list was built for
  std::vector<int64_t> lens;
  std::vector<GraphEntry> graphs;
  // Steps so far: step s uses the clip-maximum set s % 2, which the previous
  // step's phase 1 cleared, and clears the other set for the next step
  // (both sets are uploaded at -inf with the work list).
  unsigned step = 0;
  ~Batch() {
    drop_graphs();
    if (lane) {
      (void)hipStreamSynchronize(lane->stream);
      Engine::get()->release_lane(lane);
    }
  }
  void drop_graphs() {   // captured launches carry the old grid and tile list
    for (auto& g : graphs)
      for (int p = 0; p < 2; ++p) {
        if (g.exec[p]) (void)hipGraphExecDestroy(g.exec[p]);
        if (g.graph[p]) (void)hipGraphDestroy(g.graph[p]);
      }
    graphs.clear();
  }
  int grid() const {
    if (wg_per_cu <= 0) return 0;
    return Engine::get()->num_cus() * wg_per_cu;
  }
};

// Builds and uploads the tile list for `shape` (synchronous on the lane's
// stream, so no earlier step still reads the old list).
int32_t set_batch_shape(Batch* b, int shape) {
  TileLayout L;
  const bool fused = tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels,
                                 std::min(b->r.n_mfcc, b->r.n_mels), &L, shape);
  if (!fused && shape != SHAPE_DEFAULT) return ERR_INVALID;
  if (hipStreamSynchronize(b->lane->stream) != hipSuccess) return ERR_INTERNAL;
  b->drop_graphs();
  build_meta(b->r, b->lens.data(), (int)b->lens.size(), fused ? L.tile_frames : kTileFrames, &b->m);
  const size_t mb = meta_bytes(b->m);
  if (b->lane->d_meta.ensure(mb) != hipSuccess || b->lane->h_meta.ensure(mb) != hipSuccess) return ERR_ALLOC;
  b->lane->meta_key.clear();
  pack_meta(b->m, b->lane->h_meta.p);
  if (hipMemcpyAsync(b->lane->d_meta.p, b->lane->h_meta.p, mb, hipMemcpyHostToDevice, b->lane->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->lane->stream) != hipSuccess)
    return ERR_INTERNAL;
  b->mp = meta_ptrs(b->m, b->lane->d_meta.p);
  b->step = 0;
  b->shape = fused ? L.shape : SHAPE_DEFAULT;
  return prepare_batch(b->r, b->shape);   // kernel attributes, ahead of any stream capture
}


int32_t create_batch(const MMSpecParams* params, const int64_t* lengths, int32_t n_clips, Batch** out) {
  if (!lengths || n_clips <= 0 || !out) return ERR_INVALID;
  for (int i = 0; i < n_clips; ++i)
    if (lengths[i] < 0) return ERR_INVALID;
  // the engine sets the device's LDS limit, which the layout check reads
  Engine* e = Engine::get();
  if (!e) return ERR_NO_GPU;
  auto b = std::make_unique<Batch>();
  int32_t s = request_from_params(params, &b->r);
  if (s != OK) return s;
  b->lane = e->acquire_lane();
  if (!b->lane) return ERR_INTERNAL;
  s = e->tables_for(b->r, b->lane->stream, &b->tb);
  if (s != OK) return s;
  b->lens.assign(lengths, lengths + n_clips);
  // the autotuned shape for this configuration, if there is one
  int shape = e->tuned_shape(b->r);
  TileLayout L;
  if (shape != SHAPE_DEFAULT && !tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels,
                                             std::min(b->r.n_mfcc, b->r.n_mels), &L, shape))
    shape = SHAPE_DEFAULT;
  s = set_batch_shape(b.get(), shape);
  if (s != OK) return s;
  *out = b.release();
  return OK;
}

// One step with clip-maximum set `set` (0 or 1; -1: the next in turn).
int32_t run_batch(Batch* b, const void* d_signal, float* d_out, hipStream_t st, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr, int set = -1) {
  if (set < 0) set = (int)(b->step++ & 1u);
  MetaPtrs mp = b->mp;
  if (set) std::swap(mp.clip_max_key, mp.clip_max_key2);
  return enqueue_batch(Engine::get(), b->r, b->tb, b->m, mp, d_signal, d_out, &b->lane->d_scratch,
                       &b->lane->d_scratch2, b->grid(), st, e0, e1, b->shape, mp.clip_max_key2);
}

}  // namespace

extern "C" {

int32_t mm_batch_create(mm_context ctx, const MMSpecParams* params, const int64_t* lengths, int32_t n_clips,
                        mm_batch* out) try {
  (void)ctx;  // plans own their stream and workspace
  if (!out) return MM_ERR_INVALID_INPUT;
  Batch* b = nullptr;
  const int32_t s = create_batch(params, lengths, n_clips, &b);
  if (s != OK) return s;
  *out = b;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_info(mm_batch bh, int64_t* total_input_samples, int64_t* total_output_floats,
                      int64_t* total_frames, int32_t* n_tiles) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b) return MM_ERR_INVALID_INPUT;
  if (total_input_samples) *total_input_samples = b->m.total_in;
  if (total_output_floats) *total_output_floats = b->m.total_out;
  if (total_frames) {
    int64_t t = 0;
    for (int32_t f : b->m.clip_frames) t += f;
    *total_frames = t;
  }
  if (n_tiles) *n_tiles = b->m.n_tiles;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_clip(mm_batch bh, int32_t clip, int64_t* input_offset, int64_t* output_offset, int64_t* rows,
                      int64_t* frames) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || clip < 0 || clip >= b->m.n_clips) return MM_ERR_INVALID_INPUT;
  if (input_offset) *input_offset = b->m.clip_off[clip];
  if (output_offset) *output_offset = b->m.out_off[clip] * out_elem_floats(b->r);
  if (rows) *rows = out_rows(b->r);
  if (frames) *frames = b->m.clip_frames[clip];
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_run(mm_batch bh, const void* d_signal, float* d_out, void* hip_stream) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || (b->m.total_in > 0 && !d_signal) || (b->m.total_out > 0 && !d_out)) return MM_ERR_INVALID_INPUT;
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->lane->stream;
  return run_batch(b, d_signal, d_out, st);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_run_graph(mm_batch bh, const void* d_signal, float* d_out, void* hip_stream) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || (b->m.total_in > 0 && !d_signal) || (b->m.total_out > 0 && !d_out)) return MM_ERR_INVALID_INPUT;
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->lane->stream;
  // the generic (non-fused) path grows scratch buffers on the fly, which
  // stream capture does not allow: it runs eagerly
  TileLayout L;
  if (!tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels, std::min(b->r.n_mfcc, b->r.n_mels), &L,
                   b->shape))
    return run_batch(b, d_signal, d_out, st);
  const int set = (int)(b->step & 1u);
  GraphEntry* g = nullptr;
  for (auto& e : b->graphs)
    if (e.in == d_signal && e.out == d_out && e.st == st) g = &e;
  if (!g) {
    b->graphs.push_back(GraphEntry{d_signal, d_out, st, {nullptr, nullptr}, {nullptr, nullptr}});
    g = &b->graphs.back();
  }
  if (!g->exec[set]) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return MM_ERR_INTERNAL;
    const int32_t s = run_batch(b, d_signal, d_out, st, nullptr, nullptr, set);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (s != OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return s;
    }
    if (ec != hipSuccess || !graph) return MM_ERR_INTERNAL;
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return MM_ERR_INTERNAL;
    }
    g->graph[set] = graph;
    g->exec[set] = exec;
  }
  ++b->step;
  return hipGraphLaunch(g->exec[set], st) == hipSuccess ? MM_OK : MM_ERR_INTERNAL;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_set_grid(mm_batch bh, int32_t wg_per_cu) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || wg_per_cu < 0 || wg_per_cu > 64) return MM_ERR_INVALID_INPUT;
  b->wg_per_cu = wg_per_cu;
  b->drop_graphs();
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_time(mm_batch bh, const void* d_signal, float* d_out, void* hip_stream, int32_t iters,
                      float* kernel_ms, float* step_ms) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || iters <= 0) return MM_ERR_INVALID_INPUT;
  hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->lane->stream;
  std::vector<hipEvent_t> ev(4 * iters, nullptr);
  int32_t s = OK;
  for (auto& e : ev)
    if (s == OK && hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      s = ERR_INTERNAL;
    }
  for (int i = 0; i < iters && s == OK; ++i) {
    if (hipEventRecord(ev[4 * i], st) != hipSuccess) s = ERR_INTERNAL;
    if (s == OK) s = run_batch(b, d_signal, d_out, st, ev[4 * i + 1], ev[4 * i + 2]);
    if (s == OK && hipEventRecord(ev[4 * i + 3], st) != hipSuccess) s = ERR_INTERNAL;
  }
  if (s == OK && hipStreamSynchronize(st) != hipSuccess) s = ERR_INTERNAL;
  double k = 0.0, t = 0.0;
  for (int i = 0; i < iters && s == OK; ++i) {
    float a = 0.f, c = 0.f;
    if (hipEventElapsedTime(&a, ev[4 * i + 1], ev[4 * i + 2]) != hipSuccess ||
        hipEventElapsedTime(&c, ev[4 * i], ev[4 * i + 3]) != hipSuccess)
      s = ERR_INTERNAL;
    k += a;
    t += c;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (s != OK) return s;
  if (kernel_ms) *kernel_ms = (float)(k / iters);
  if (step_ms) *step_ms = (float)(t / iters);
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_autotune(mm_batch bh, const void* d_signal, float* d_out, int32_t iters, int32_t* best_wg_per_cu,
                          float* best_ms) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || iters <= 0) return MM_ERR_INVALID_INPUT;
  // Candidates: every workgroup shape the fused kernel offers for this FFT
  // size (tile frames and waves per workgroup), each with persistent grids
  // of 1..8 workgroups per CU, timed as whole steps on this batch. Replaces
  // fixed per-device thresholds with a measurement on the device at hand.
  int shapes[4];
  int ns = tile_shapes(b->r.n_fft, shapes);
  TileLayout L;
  if (!tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels, std::min(b->r.n_mfcc, b->r.n_mels), &L))
    ns = 0;   // generic path: only the grid matters (and it is not used)
  if (ns == 0) {
    shapes[0] = SHAPE_DEFAULT;
    ns = 1;
  }
  const int cands[] = {1, 2, 3, 4, 6, 8};
  const int saved_wg = b->wg_per_cu, saved_shape = b->shape;
  int best = 0, best_shape = SHAPE_DEFAULT;
  float best_t = 1e30f;
  for (int si = 0; si < ns; ++si) {
    if (!tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels, std::min(b->r.n_mfcc, b->r.n_mels),
                     &L, shapes[si]) &&
        shapes[si] != SHAPE_DEFAULT)
      continue;   // this kind or precision does not fit the shape's LDS layout
    int32_t s = set_batch_shape(b, shapes[si]);
    for (int c : cands) {
      if (s != OK) break;
      b->wg_per_cu = c;
      float k = 0.f, t = 0.f;
      s = mm_batch_time(b, d_signal, d_out, nullptr, 1, &k, &t);  // warm-up
      if (s == OK) s = mm_batch_time(b, d_signal, d_out, nullptr, iters, &k, &t);
      if (s == OK && t < best_t) {
        best_t = t;
        best = c;
        best_shape = b->shape;
      }
    }
    if (s != OK) {
      b->wg_per_cu = saved_wg;
      (void)set_batch_shape(b, saved_shape);
      return s;
    }
  }
  int32_t s = set_batch_shape(b, best_shape);
  if (s != OK) return s;
  mm_batch_set_grid(b, best);
  Engine::get()->record_tuning(b->r, best_shape, best);
  if (best_wg_per_cu) *best_wg_per_cu = best;
  if (best_ms) *best_ms = best_t;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_set_shape(mm_batch bh, int32_t shape) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b || shape < 0) return MM_ERR_INVALID_INPUT;
  if (shape == b->shape) return MM_OK;
  return set_batch_shape(b, shape);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_batch_launch_shape(mm_batch bh, int32_t* shape, int32_t* waves_per_wg, int32_t* tile_frames,
                              int32_t* grid, int32_t* tiles_per_wg) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b) return MM_ERR_INVALID_INPUT;
  TileLayout L;
  const bool fused = tile_layout(b->r.n_fft, b->r.out_kind, b->r.precision, b->r.n_mels,
                                 std::min(b->r.n_mfcc, b->r.n_mels), &L, b->shape);
  int tpw = 0, g = 0;
  if (fused) g = launch_grid(Engine::get(), b->r, b->m.n_tiles, L, b->grid(), &tpw);
  if (shape) *shape = fused ? L.shape : SHAPE_DEFAULT;
  if (waves_per_wg) *waves_per_wg = fused ? L.nw : 0;
  if (tile_frames) *tile_frames = fused ? L.tile_frames : 0;
  if (grid) *grid = g;
  if (tiles_per_wg) *tiles_per_wg = tpw;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

#ifdef MM_PHASE_TIMING
// Instrumented builds only: device buffer for per-wave phase cycle sums.
int32_t mm_debug_set_phase_buffer(void* d_buf) try {
  set_phase_cycles_buffer(static_cast<long long*>(d_buf));
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}
#endif

int32_t mm_batch_destroy(mm_batch bh) try {
  Batch* b = static_cast<Batch*>(bh);
  if (!b) return MM_ERR_INVALID_INPUT;
  delete b;
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_spectral_batch(mm_context ctx, const MMSpecParams* params, const float* const* clips,
                          const int64_t* lengths, int32_t n_clips, MMBuffer* outs) try {
  if (!ctx || !params || !clips || !lengths || n_clips <= 0 || !outs) return MM_ERR_INVALID_INPUT;
  if (params->input_dtype != MM_INPUT_F32) return MM_ERR_INVALID_INPUT;
  for (int i = 0; i < n_clips; ++i)
    if (!clips[i] || lengths[i] <= 0) return MM_ERR_INVALID_INPUT;
  Batch* b = nullptr;
  int32_t s = create_batch(params, lengths, n_clips, &b);
  if (s != OK) return s;
  std::unique_ptr<Batch> hold(b);
  Lane* L = b->lane;
  const size_t in_bytes = (size_t)b->m.total_in * 4, out_bytes = (size_t)b->m.total_out * 4;
  if (L->h_in.ensure(in_bytes) != hipSuccess || L->d_in.ensure(in_bytes) != hipSuccess ||
      L->d_out.ensure(out_bytes) != hipSuccess || L->h_out.ensure(out_bytes) != hipSuccess)
    return MM_ERR_ALLOC_FAILED;
  for (int i = 0; i < n_clips; ++i)
    std::memcpy(static_cast<float*>(L->h_in.p) + b->m.clip_off[i], clips[i], (size_t)lengths[i] * 4);
  if (hipMemcpyAsync(L->d_in.p, L->h_in.p, in_bytes, hipMemcpyHostToDevice, L->stream) != hipSuccess)
    return MM_ERR_INTERNAL;
  s = run_batch(b, L->d_in.p, static_cast<float*>(L->d_out.p), L->stream);
  if (s != OK) {
    (void)hipStreamSynchronize(L->stream);
    return s;
  }
  if (hipMemcpyAsync(L->h_out.p, L->d_out.p, out_bytes, hipMemcpyDeviceToHost, L->stream) != hipSuccess ||
      hipStreamSynchronize(L->stream) != hipSuccess)
    return MM_ERR_INTERNAL;
  const int ef = out_elem_floats(b->r);
  const int64_t rows = out_rows(b->r);
  for (int i = 0; i < n_clips; ++i) {
    const int64_t T = b->m.clip_frames[i];
    const int64_t cnt = rows * T;
    MMBuffer* o = &outs[i];
    std::memset(o, 0, sizeof(*o));
    o->ndim = 2;
    o->shape[0] = rows;
    o->shape[1] = T;
    o->dtype = (ef == 2) ? 1 : 0;
    o->count = cnt;
    if (cnt > 0) {
      o->data = static_cast<float*>(std::malloc((size_t)cnt * ef * 4));
      if (!o->data) {
        for (int j = 0; j <= i; ++j) mm_buffer_free(&outs[j]);
        return MM_ERR_ALLOC_FAILED;
      }
      std::memcpy(o->data, static_cast<float*>(L->h_out.p) + b->m.out_off[i] * ef, (size_t)cnt * ef * 4);
    }
  }
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

}  // extern "C"
