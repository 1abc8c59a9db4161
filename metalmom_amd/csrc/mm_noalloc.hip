// mm_spectral_device_noalloc: one clip on device memory, through the fused
// tile kernel, with no device allocation. The clip's work list and the
// top_db scratch (clip maximum, per-tile minima) live in a file-scope device
// array of fixed size, written by a small kernel on the caller's stream, so
// the call is a sequence of kernel launches only (capturable in a hipGraph).
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (different streams or threads without ordering) race on it.
#include <mutex>

#include "../../include/metalmom.h"
#include "mm_engine.h"

namespace mm {

// Largest work list the array holds: 1 MiB (meta_bytes of one clip), about
// 37,000 tiles of 16 frames.
constexpr size_t kNoallocScratchBytes = 1u << 20;

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

// Device address of `scratch` on each device, from hipGetSymbolAddress with
// that device current (the host-side name is not a device address).
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int32_t scratch_ptr(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return ERR_INTERNAL;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ERR_INTERNAL;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return OK;
}

// Work list of one clip (the same records build_meta writes on the host):
// tile t = frames [t * tile_frames, ...) of clip 0, its prefetch record, and
// the clip entries.
__global__ void k_single_meta(int64_t len, int T, int n_fft, int hop, int center, int tile_frames, int n_tiles,
                              int64_t* clip_off, int64_t* clip_len, int64_t* out_off, int32_t* clip_frames,
                              int32_t* tile_clip, int32_t* tile_f0, TilePf* tile_pf) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    clip_off[0] = 0;
    clip_len[0] = len;
    out_off[0] = 0;
    clip_frames[0] = T;
  }
  if (t < n_tiles) {
    const int f = t * tile_frames;
    tile_clip[t] = 0;
    tile_f0[t] = f;
    const int64_t pad = center ? n_fft / 2 : 0;
    TilePf p;
    p.base = (long long)((int64_t)f * hop - pad);
    p.mask = 0u;
    p.pad = 0;
    for (int w = 0; w < tile_frames && w < 32; ++w) {
      const int64_t s0 = (int64_t)(f + w) * hop - pad;
      if (f + w < T && s0 >= 0 && s0 + n_fft <= len && (s0 & 1) == 0) p.mask |= 1u << w;
    }
    tile_pf[t] = p;
  }
}

}  // namespace mm

using namespace mm;

extern "C" int32_t mm_spectral_device_noalloc(const MMSpecParams* params, const float* d_signal,
                                              int64_t signal_length, float* d_out, int64_t* rows_out,
                                              int64_t* frames_out, void* stream) try {
  if (!params || signal_length <= 0 || !d_signal) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  int32_t s = request_from_params(params, &r);
  if (s != OK) return s;
  if (r.in_dtype != IN_F32) return MM_ERR_INVALID_INPUT;
  TileLayout L;
  if (!tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, std::min(r.n_mfcc, r.n_mels), &L))
    return MM_ERR_INVALID_INPUT;   // outside the fused kernel: the generic path needs scratch of its own
  const int64_t T = frame_count(signal_length, r.n_fft, r.hop, r.center);
  if (rows_out) *rows_out = out_rows(r);
  if (frames_out) *frames_out = T;
  if (T == 0) return MM_OK;
  if (!d_out) return MM_ERR_INVALID_INPUT;
  BatchMeta m;
  m.n_clips = 1;
  m.n_tiles = (int)((T + L.tile_frames - 1) / L.tile_frames);
  m.total_in = signal_length;
  m.total_out = out_rows(r) * T * out_elem_floats(r);
  if (T > (int64_t)1 << 30 || meta_bytes(m) > kNoallocScratchBytes) return MM_ERR_INVALID_INPUT;   // too long
  Engine* eng = Engine::get();
  if (!eng) return MM_ERR_METAL_UNAVAILABLE;
  const Tables* tb = eng->tables_resident(r);
  if (!tb) return MM_ERR_ALLOC_FAILED;   // tables not built yet (that needs an allocating call first)
  void* base = nullptr;
  s = scratch_ptr(&base);
  if (s != OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const MetaPtrs mp = meta_ptrs(m, base);
  hipLaunchKernelGGL(k_single_meta, dim3((m.n_tiles + 255) / 256), dim3(256), 0, st, signal_length, (int)T,
                     r.n_fft, r.hop, r.center, L.tile_frames, m.n_tiles, const_cast<int64_t*>(mp.clip_off),
                     const_cast<int64_t*>(mp.clip_len), const_cast<int64_t*>(mp.out_off),
                     const_cast<int32_t*>(mp.clip_frames), const_cast<int32_t*>(mp.tile_clip),
                     const_cast<int32_t*>(mp.tile_f0), const_cast<TilePf*>(mp.tile_pf));
  if (hipGetLastError() != hipSuccess) return MM_ERR_INTERNAL;
  // the default persistent grid of the layout (no tuning-file lookup)
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, lds_limit_bytes() / std::max<size_t>(L.lds_bytes, 1)));
  return enqueue_batch(eng, r, tb, m, mp, d_signal, d_out, nullptr, nullptr, eng->num_cus() * per_cu, st);
} catch (...) {   // no exception leaves the C ABI
  return MM_ERR_INTERNAL;
}
