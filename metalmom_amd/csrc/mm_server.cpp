// Request server with continuous batching (additive exports mm_server_*).
//
// Clients submit single clips of any request type from any thread. A
// scheduler thread admits queued requests into the next device step while the
// HBM budget and a free lane allow it, groups the step's requests by
// configuration (one persistent fused launch per group), and keeps up to
// kLanes steps in flight on separate streams, so requests that arrive while
// others run on the GPU join the next step and overlap with them (H2D of step
// n+1 and D2H of step n-1 run next to the kernels of step n). Device state
// (tables, work lists, lane buffers) stays resident between steps.
//
// Host copies stay off the scheduler thread: submit() copies the clip straight
// into a pinned input ring (one copy, on the client's thread), from which the
// step's H2D copies read; each step's D2H lands in a pinned output ring, and
// the client's poll()/wait() copies its own result out. When a ring is full the
// request falls back to heap memory and the lane's staging buffers.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/metalmom.h"
#include "mm_engine.h"

using namespace mm;

namespace {

constexpr int kLanes = 3;

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// FIFO allocations out of one pinned block. A range becomes reusable once it
// and every range allocated before it have been released.
class PinnedRing {
 public:
  ~PinnedRing() {
    if (base_) (void)hipHostFree(base_);
  }
  bool init(size_t cap) {
    if (hipHostMalloc(reinterpret_cast<void**>(&base_), cap, hipHostMallocDefault) != hipSuccess) {
      base_ = nullptr;
      return false;
    }
    cap_ = cap;
    return true;
  }
  char* alloc(size_t n) {
    n = (n + 255) & ~(size_t)255;
    std::lock_guard<std::mutex> g(mu_);
    if (!base_ || n == 0 || n > cap_) return nullptr;
    size_t off;
    if (live_.empty()) {
      off = 0;
    } else {
      const size_t tail = live_.front().off, head = live_.back().off + live_.back().len;
      if (head > tail) {  // live range [tail, head): room at the end, else at the start
        if (cap_ - head >= n) off = head;
        else if (tail >= n) off = 0;
        else return nullptr;
      } else {  // wrapped: live ranges [tail, cap) and [0, head)
        if (tail - head >= n) off = head;
        else return nullptr;
      }
    }
    live_.push_back({off, n, false});
    return base_ + off;
  }
  void release(const char* p) {
    std::lock_guard<std::mutex> g(mu_);
    const size_t off = (size_t)(p - base_);
    for (auto& b : live_)
      if (b.off == off && !b.freed) {
        b.freed = true;
        break;
      }
    while (!live_.empty() && live_.front().freed) live_.pop_front();
  }

 private:
  struct Range {
    size_t off, len;
    bool freed;
  };
  std::mutex mu_;
  char* base_ = nullptr;
  size_t cap_ = 0;
  std::deque<Range> live_;
};

// One step's results in the output ring, released when the last of its
// requests has been collected.
struct OutBlock {
  char* p = nullptr;
  int refs = 0;
};

struct Req {
  int64_t id = 0;
  SpecRequest r;
  int64_t len = 0;             // samples the analysis sees (after PCM conversion and resampling)
  int fmt = 0;                 // 0: float32 samples; else the PcmFormat of interleaved PCM
  int channels = 1, mix = 0, src_sr = 0;
  int64_t frames = 0;          // PCM frames (fmt != 0)
  size_t data_bytes = 0;       // bytes as submitted
  const char* pin = nullptr;   // submitted data in the input ring, or
  std::vector<char> heap;      // on the heap when the ring was full
  const char* res = nullptr;   // result in the output ring (block blk), or `out`
  OutBlock* blk = nullptr;
  double t_submit = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 on device, 2 done
  int32_t status = OK;
  float* out = nullptr;
  int64_t rows = 0, cols = 0;
  size_t dev_bytes = 0;
};

struct Group {
  SpecRequest r;
  std::vector<Req*> reqs;
  BatchMeta m;
  size_t in_off = 0, out_off = 0, meta_off = 0;  // byte offsets in the lane buffers
};

struct Step {
  Lane* lane = nullptr;
  hipEvent_t done = nullptr;
  OutBlock* blk = nullptr;  // null: results staged in the lane's h_out
  std::vector<Group> groups;
  size_t in_bytes = 0, out_bytes = 0, meta_bytes_ = 0, dev_bytes = 0;
  int32_t status = OK;
};

bool same_plan(const SpecRequest& a, const SpecRequest& b) {
  return a.out_kind == b.out_kind && a.sr == b.sr && a.n_fft == b.n_fft && a.hop == b.hop && a.win == b.win &&
         a.center == b.center && a.pad_mode == b.pad_mode && a.power == b.power && a.n_mels == b.n_mels &&
         a.fmin == b.fmin && a.fmax == b.fmax && a.n_mfcc == b.n_mfcc && a.top_db == b.top_db &&
         a.precision == b.precision && a.in_dtype == b.in_dtype && a.fb_kind == b.fb_kind && a.tuning == b.tuning;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes a request holds while in a step: its input and output, its
// share of the work list (per-clip records and at most T / 8 + 1 tiles of
// 48 bytes), plus for configurations outside the fused kernel the generic
// path's complex spectrogram scratch (T x F x 8 bytes) and, for MFCC, its
// mel scratch.
size_t request_bytes(const SpecRequest& r, int64_t len) {
  const int64_t T = frame_count(len, r.n_fft, r.hop, r.center);
  size_t b = a256((size_t)len * 4) + a256((size_t)(out_rows(r) * T * out_elem_floats(r)) * 4) + 1024 +
             (size_t)(T / 8 + 1) * 48;
  TileLayout L;
  if (!tile_layout(r.n_fft, r.out_kind, r.precision, r.n_mels, std::min(r.n_mfcc, r.n_mels), &L)) {
    b += a256((size_t)T * (size_t)(r.n_fft / 2 + 1) * 8);
    if (r.out_kind == OUT_MFCC) b += a256((size_t)r.n_mels * (size_t)T * 4 + 256);
  }
  return b;
}

// Device bytes a lane's buffers hold, and their release (the lane is idle).
size_t lane_dev_bytes(const Lane* L) {
  return L->d_in.n + L->d_out.n + L->d_meta.n + L->d_scratch.n + L->d_scratch2.n + L->d_small.n;
}
void release_dev(Lane* L) {
  L->d_in.release();
  L->d_out.release();
  L->d_meta.release();
  L->d_scratch.release();
  L->d_scratch2.release();
  L->d_small.release();
  L->meta_key.clear();
}

// Device staging of a PCM request: its raw bytes, and its converted samples
// when they still have to be resampled.
size_t pcm_stage_bytes(const Req& q) {
  if (!q.fmt) return 0;
  return a256(q.data_bytes) + (q.src_sr != q.r.sr ? a256((size_t)q.frames * 4) : 0);
}

class Server {
 public:
  Server(size_t budget, int max_batch) : budget_(budget), max_batch_(std::max(1, max_batch)) {}
  ~Server() { stop(); }

  int32_t start() {
    Engine* e = Engine::get();
    if (!e) return ERR_NO_GPU;
    // a quarter of the HBM budget each, 32..512 MiB; a ring that cannot be
    // pinned leaves every request on the heap path
    const size_t ring = std::min<size_t>(std::max<size_t>(budget_ / 4, (size_t)32 << 20), (size_t)512 << 20);
    (void)in_ring_.init(ring);
    (void)out_ring_.init(ring);
    for (int i = 0; i < kLanes; ++i) {
      Lane* l = e->acquire_lane();
      if (!l) return ERR_INTERNAL;
      release_dev(l);   // pooled lanes may hold buffers of earlier users
      free_lanes_.push_back(l);
      lanes_.push_back(l);
    }
    th_ = std::thread([this] { loop(); });
    return OK;
  }

  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (stopping_) return;
      stopping_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
    for (auto& s : inflight_) {
      (void)hipEventSynchronize(s.done);
      (void)hipEventDestroy(s.done);
      delete s.blk;
      free_lanes_.push_back(s.lane);
    }
    inflight_.clear();
    for (Lane* l : free_lanes_) Engine::get()->release_lane(l);
    free_lanes_.clear();
    for (auto& kv : reqs_) {
      if (kv.second->out) std::free(kv.second->out);
      drop_result(kv.second);
      delete kv.second;
    }
    reqs_.clear();
  }

  // A float32 clip of len samples, or (fmt != 0) `frames` frames of
  // interleaved PCM at src_sr that become a clip of len samples at r.sr.
  int64_t submit(const SpecRequest& r, const void* data, int64_t len, int fmt = 0, int64_t frames = 0,
                 int channels = 1, int mix = 0, int src_sr = 0) {
    auto* q = new Req();
    q->r = r;
    q->len = len;
    q->fmt = fmt;
    q->frames = frames;
    q->channels = channels;
    q->mix = mix;
    q->src_sr = src_sr;
    q->data_bytes = fmt ? (size_t)frames * channels * pcm_bytes(fmt) : (size_t)len * 4;
    char* pin = in_ring_.alloc(q->data_bytes);
    if (pin) {
      std::memcpy(pin, data, q->data_bytes);
      q->pin = pin;
    } else {
      const char* b = static_cast<const char*>(data);
      q->heap.assign(b, b + q->data_bytes);
    }
    q->dev_bytes = request_bytes(r, len) + pcm_stage_bytes(*q);
    q->t_submit = now_s();
    std::lock_guard<std::mutex> g(mu_);
    q->id = next_id_++;
    if (q->dev_bytes > budget_) {  // can never be admitted
      q->status = ERR_ALLOC;
      q->state = 2;
      q->t_done = q->t_submit;
      release_input(q);
    } else {
      queue_.push_back(q);
    }
    reqs_[q->id] = q;
    ++submitted_;
    cv_.notify_all();
    return q->id;
  }

  // 1: finished (out filled, request forgotten), 0: pending, <0: error/unknown
  int32_t poll(int64_t id, MMBuffer* out, double* latency, bool wait) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return ERR_INVALID;
    if (wait) {
      // another thread may collect the same id meanwhile: look it up again
      done_cv_.wait(g, [&] {
        it = reqs_.find(id);
        return it == reqs_.end() || it->second->state == 2;
      });
      if (it == reqs_.end()) return ERR_INVALID;
    }
    Req* q = it->second;
    if (q->state != 2) return 0;
    reqs_.erase(it);
    g.unlock();
    done_cv_.notify_all();   // a second waiter on this id returns ERR_INVALID
    if (latency) *latency = q->t_done - q->t_submit;
    int32_t s = q->status;
    if (s == OK && q->res) {  // copy the result out of the output ring
      const size_t n = (size_t)(q->rows * q->cols * out_elem_floats(q->r)) * 4;
      if (out) {
        q->out = static_cast<float*>(std::malloc(n));
        if (q->out) std::memcpy(q->out, q->res, n);
        else s = ERR_ALLOC;
      }
      drop_result(q);
    }
    if (s == OK && out) {
      std::memset(out, 0, sizeof(*out));
      out->ndim = 2;
      out->shape[0] = q->rows;
      out->shape[1] = q->cols;
      out->dtype = (q->r.out_kind == OUT_STFT_COMPLEX) ? 1 : 0;
      out->count = q->rows * q->cols;
      out->data = q->out;
      q->out = nullptr;
    }
    if (q->out) std::free(q->out);
    delete q;
    return s == OK ? 1 : s;
  }

  // Device bytes the server's lanes hold (inputs, outputs, work lists,
  // scratch). A lane keeps its buffers after a step only while they hold at
  // most budget / 3 bytes, and a step's buffers hold at most 1.5 times its
  // own bytes, which the admission caps at the budget over all steps in
  // flight: at most 2.5 x the budget in all, at most the budget when idle.
  size_t device_bytes() const { return held_.load(); }

  void stats(int64_t* submitted, int64_t* completed, int64_t* steps, int64_t* frames) {
    std::lock_guard<std::mutex> g(mu_);
    if (submitted) *submitted = submitted_;
    if (completed) *completed = completed_;
    if (steps) *steps = steps_;
    if (frames) *frames = frames_;
  }

 private:
  void loop() {
    while (true) {
      retire(false);
      std::vector<Req*> take;
      {
        std::unique_lock<std::mutex> g(mu_);
        if (stopping_ && queue_.empty()) break;
        if (queue_.empty() || free_lanes_.empty()) {
          if (inflight_.empty()) {
            cv_.wait(g, [&] { return stopping_ || !queue_.empty(); });
            continue;
          }
          g.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(20));
          continue;
        }
        // admission: FIFO while the HBM budget allows
        size_t step_bytes = 0;
        while (!queue_.empty() && (int)take.size() < max_batch_) {
          Req* q = queue_.front();
          if (in_use_ + step_bytes + q->dev_bytes > budget_) break;
          step_bytes += q->dev_bytes;
          q->state = 1;
          take.push_back(q);
          queue_.pop_front();
        }
        if (take.empty()) {  // budget full: wait for a step to retire
          g.unlock();
          retire(true);
          continue;
        }
        in_use_ += step_bytes;
      }
      launch(take);
    }
    while (!inflight_.empty()) retire(true);
  }

  void launch(std::vector<Req*>& take) {
    Step s;
    {
      std::lock_guard<std::mutex> g(mu_);
      s.lane = free_lanes_.back();
      free_lanes_.pop_back();
    }
    try {
      prepare(s, take);
    } catch (...) {   // an allocation failed: nothing of the step may stay in flight
      (void)hipStreamSynchronize(s.lane->stream);
      if (s.blk) {
        out_ring_.release(s.blk->p);
        delete s.blk;
      }
      if (s.done) (void)hipEventDestroy(s.done);
      fail(take, s.lane);
      return;
    }
    std::lock_guard<std::mutex> g(mu_);
    ++steps_;
    inflight_.push_back(std::move(s));
  }

  // Ends the requests of a step that could not be launched with ERR_ALLOC and
  // returns its lane.
  void fail(std::vector<Req*>& take, Lane* lane) {
    const double t = now_s();
    size_t bytes = 0;
    for (Req* q : take) {
      release_input(q);
      bytes += q->dev_bytes;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      for (Req* q : take) {
        q->status = ERR_ALLOC;
        q->t_done = t;
        q->state = 2;
        ++completed_;
      }
      in_use_ -= bytes;
      free_lanes_.push_back(lane);
    }
    done_cv_.notify_all();
    cv_.notify_all();
  }

  // Lays out the step, stages and enqueues its copies and kernels on the
  // lane's stream and records its completion event (s.done).
  void prepare(Step& s, std::vector<Req*>& take) {
    Engine* eng = Engine::get();
    for (Req* q : take) {
      s.dev_bytes += q->dev_bytes;
      Group* grp = nullptr;
      for (auto& G : s.groups)
        if (same_plan(G.r, q->r)) grp = &G;
      if (!grp) {
        s.groups.push_back(Group());
        grp = &s.groups.back();
        grp->r = q->r;
      }
      grp->reqs.push_back(q);
    }
    Lane* L = s.lane;
    hipStream_t st = L->stream;
    // layout of inputs, outputs and work lists in the lane buffers
    size_t heap_bytes = 0, stage_bytes = 0;
    for (auto& G : s.groups) {
      std::vector<int64_t> lens;
      for (Req* q : G.reqs) lens.push_back(q->len);
      TileLayout TL;
      const bool fused =
          tile_layout(G.r.n_fft, G.r.out_kind, G.r.precision, G.r.n_mels, std::min(G.r.n_mfcc, G.r.n_mels), &TL);
      build_meta(G.r, lens.data(), (int)lens.size(), fused ? TL.tile_frames : kTileFrames, &G.m);
      G.in_off = s.in_bytes;
      s.in_bytes += a256((size_t)G.m.total_in * 4);
      G.out_off = s.out_bytes;
      s.out_bytes += a256((size_t)G.m.total_out * 4);
      G.meta_off = s.meta_bytes_;
      s.meta_bytes_ += a256(meta_bytes(G.m));
      for (const Req* q : G.reqs) {
        if (!q->pin) heap_bytes += a256(q->data_bytes);
        stage_bytes += pcm_stage_bytes(*q);
      }
    }
    int32_t st_code = OK;
    auto chk = [&](hipError_t e) {
      if (e != hipSuccess && st_code == OK) st_code = ERR_INTERNAL;
    };
    // lane staging: work lists, then the clips that are not in the input ring
    chk(L->h_in.ensure(s.meta_bytes_ + heap_bytes));
    chk(L->d_in.ensure(s.in_bytes + stage_bytes));   // clips | PCM staging
    chk(L->d_meta.ensure(s.meta_bytes_));
    L->meta_key.clear();
    chk(L->d_out.ensure(s.out_bytes));
    update_held();
    if (s.out_bytes > 0) {
      if (char* p = out_ring_.alloc(s.out_bytes)) {
        s.blk = new OutBlock();
        s.blk->p = p;
      }
    }
    if (!s.blk) chk(L->h_out.ensure(s.out_bytes));
    if (st_code == OK) {
      char* hin = static_cast<char*>(L->h_in.p);
      char* heap_stage = hin + s.meta_bytes_;
      char* din = static_cast<char*>(L->d_in.p);
      char* dstage = din + s.in_bytes;
      for (auto& G : s.groups) {
        for (size_t i = 0; i < G.reqs.size(); ++i) {
          const Req* q = G.reqs[i];
          char* dst = din + G.in_off + (size_t)G.m.clip_off[i] * 4;
          const void* src = q->pin;
          if (!src) {  // heap data: stage it in the lane's pinned buffer
            std::memcpy(heap_stage, q->heap.data(), q->data_bytes);
            src = heap_stage;
            heap_stage += a256(q->data_bytes);
          }
          if (!q->fmt) {
            chk(hipMemcpyAsync(dst, src, q->data_bytes, hipMemcpyHostToDevice, st));
            continue;
          }
          // PCM: raw bytes -> one float channel (-> resampled) into the clip's slot
          char* raw = dstage;
          dstage += a256(q->data_bytes);
          chk(hipMemcpyAsync(raw, src, q->data_bytes, hipMemcpyHostToDevice, st));
          if (q->src_sr == G.r.sr) {
            chk(launch_pcm_to_mono(raw, q->frames, q->channels, q->fmt, q->mix, reinterpret_cast<float*>(dst),
                                   eng->num_cus(), st));
          } else {
            float* mono = reinterpret_cast<float*>(dstage);
            dstage += a256((size_t)q->frames * 4);
            chk(launch_pcm_to_mono(raw, q->frames, q->channels, q->fmt, q->mix, mono, eng->num_cus(), st));
            if (enqueue_resample(mono, q->frames, q->src_sr, G.r.sr, reinterpret_cast<float*>(dst), st) != OK &&
                st_code == OK)
              st_code = ERR_INTERNAL;
          }
        }
        pack_meta(G.m, hin + G.meta_off);
      }
      chk(hipMemcpyAsync(L->d_meta.p, hin, s.meta_bytes_, hipMemcpyHostToDevice, st));
      for (auto& G : s.groups) {
        if (st_code != OK) break;
        const Tables* tb = nullptr;
        int32_t r = eng->tables_for(G.r, st, &tb);
        if (r == OK) {
          const MetaPtrs mp = meta_ptrs(G.m, static_cast<char*>(L->d_meta.p) + G.meta_off);
          r = enqueue_batch(eng, G.r, tb, G.m, mp, static_cast<char*>(L->d_in.p) + G.in_off,
                            static_cast<char*>(L->d_out.p) + G.out_off, &L->d_scratch, &L->d_scratch2, 0, st);
        }
        if (r != OK) st_code = r;
      }
      chk(hipMemcpyAsync(s.blk ? s.blk->p : L->h_out.p, L->d_out.p, s.out_bytes, hipMemcpyDeviceToHost, st));
    }
    chk(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    if (s.done) chk(hipEventRecord(s.done, st));
    s.status = st_code;
  }

  void release_input(Req* q) {
    if (q->pin) in_ring_.release(q->pin);
    q->pin = nullptr;
    std::vector<char>().swap(q->heap);
  }

  // Lets go of q's share of its step's output block.
  void drop_result(Req* q) {
    OutBlock* b = q->blk;
    q->blk = nullptr;
    q->res = nullptr;
    if (!b) return;
    bool last;
    {
      std::lock_guard<std::mutex> g(blk_mu_);
      last = --b->refs == 0;
    }
    if (last) {
      out_ring_.release(b->p);
      delete b;
    }
  }

  // Completes finished steps (the oldest first). block: wait for the oldest.
  void retire(bool block) {
    while (true) {
      Step* s = nullptr;
      {
        std::lock_guard<std::mutex> g(mu_);
        if (inflight_.empty()) return;
        s = &inflight_.front();
      }
      if (block) {
        (void)hipEventSynchronize(s->done);
        block = false;
      } else if (hipEventQuery(s->done) == hipErrorNotReady) {
        return;
      }
      const char* hout = s->blk ? s->blk->p : static_cast<const char*>(s->lane->h_out.p);
      const double t = now_s();
      int64_t frames = 0;
      int refs = 0;
      for (auto& G : s->groups) {
        const int ef = out_elem_floats(G.r);
        const int64_t rows = out_rows(G.r);
        for (size_t i = 0; i < G.reqs.size(); ++i) {
          Req* q = G.reqs[i];
          q->rows = rows;
          q->cols = G.m.clip_frames[i];
          frames += q->cols;
          q->status = s->status;
          const size_t n = (size_t)(rows * q->cols * ef) * 4;
          const char* src = hout + G.out_off + (size_t)G.m.out_off[i] * ef * 4;
          if (q->status == OK && n > 0) {
            if (s->blk) {  // the client copies it out of the ring
              q->res = src;
              q->blk = s->blk;
              ++refs;
            } else {
              q->out = static_cast<float*>(std::malloc(n));
              if (!q->out) q->status = ERR_ALLOC;
              else std::memcpy(q->out, src, n);
            }
          }
          release_input(q);
        }
      }
      if (s->blk) {
        bool unused;
        {
          std::lock_guard<std::mutex> g(blk_mu_);
          s->blk->refs += refs;
          unused = s->blk->refs == 0;
        }
        if (unused) {
          out_ring_.release(s->blk->p);
          delete s->blk;
        }
        s->blk = nullptr;
      }
      (void)hipEventDestroy(s->done);
      // the idle lane keeps its buffers only up to a third of the budget
      if (lane_dev_bytes(s->lane) > budget_ / kLanes) {
        release_dev(s->lane);
        update_held();
      }
      {
        std::lock_guard<std::mutex> g(mu_);
        for (auto& G : s->groups)
          for (Req* q : G.reqs) {
            q->t_done = t;
            q->state = 2;
            ++completed_;
          }
        frames_ += frames;
        in_use_ -= s->dev_bytes;
        free_lanes_.push_back(s->lane);
        inflight_.pop_front();
      }
      done_cv_.notify_all();
      cv_.notify_all();
    }
  }

  // held_ = device bytes of all lanes; written by the scheduler thread, the
  // only one that resizes lane buffers
  void update_held() {
    size_t b = 0;
    for (const Lane* l : lanes_) b += lane_dev_bytes(l);
    held_.store(b);
  }

  size_t budget_;
  int max_batch_;
  std::vector<Lane*> lanes_;
  std::atomic<size_t> held_{0};
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Req*> queue_;
  std::unordered_map<int64_t, Req*> reqs_;
  std::deque<Step> inflight_;
  std::vector<Lane*> free_lanes_;
  size_t in_use_ = 0;
  PinnedRing in_ring_, out_ring_;
  std::mutex blk_mu_;
  int64_t next_id_ = 1, submitted_ = 0, completed_ = 0, steps_ = 0, frames_ = 0;
  bool stopping_ = false;
  std::thread th_;
};


}  // namespace

extern "C" {

int32_t mm_server_create(mm_context ctx, int64_t hbm_budget_bytes, int32_t max_batch, mm_server* out) try {
  (void)ctx;
  if (!out || hbm_budget_bytes <= 0) return MM_ERR_INVALID_INPUT;
  std::unique_ptr<Server> s(new Server((size_t)hbm_budget_bytes, max_batch > 0 ? max_batch : 256));
  const int32_t r = s->start();   // may throw (thread creation): s is freed either way
  if (r != OK) return r;
  *out = s.release();
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int64_t mm_server_submit(mm_server srv, const MMSpecParams* params, const float* signal_data, int64_t signal_length) {
  if (!srv || !signal_data || signal_length <= 0) return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  int32_t s = request_from_params(params, &r);
  if (s == OK && r.in_dtype != IN_F32) s = ERR_INVALID;   // host clips are float32
  if (s != OK) return s;
  try {
    return static_cast<Server*>(srv)->submit(r, signal_data, signal_length);
  } catch (const std::bad_alloc&) {   // no exception leaves the C ABI
    return MM_ERR_ALLOC_FAILED;
  }
}

int64_t mm_server_submit_pcm(mm_server srv, const MMSpecParams* params, const void* pcm, int64_t frames,
                             int32_t channels, int32_t pcm_format, int32_t source_sr, int32_t mono) {
  if (!srv || !pcm || frames <= 0 || channels < 1 || channels > 64 || !pcm_valid(pcm_format) || source_sr <= 0)
    return MM_ERR_INVALID_INPUT;
  SpecRequest r;
  int32_t s = request_from_params(params, &r);
  if (s == OK && r.in_dtype != IN_F32) s = ERR_INVALID;   // the analysis input is float32
  if (s != OK) return s;
  const int64_t len = resample_length(frames, source_sr, r.sr);
  if (len <= 0) return MM_ERR_INVALID_INPUT;
  try {
    return static_cast<Server*>(srv)->submit(r, pcm, len, pcm_format, frames, channels,
                                             (mono != 0 && channels > 1) ? 1 : 0, source_sr);
  } catch (const std::bad_alloc&) {   // no exception leaves the C ABI
    return MM_ERR_ALLOC_FAILED;
  }
}

int32_t mm_server_poll(mm_server srv, int64_t request_id, MMBuffer* out, double* latency_s) try {
  if (!srv) return MM_ERR_INVALID_INPUT;
  return static_cast<Server*>(srv)->poll(request_id, out, latency_s, false);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_server_wait(mm_server srv, int64_t request_id, MMBuffer* out, double* latency_s) try {
  if (!srv) return MM_ERR_INVALID_INPUT;
  return static_cast<Server*>(srv)->poll(request_id, out, latency_s, true);
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_server_stats(mm_server srv, int64_t* submitted, int64_t* completed, int64_t* steps, int64_t* frames) try {
  if (!srv) return MM_ERR_INVALID_INPUT;
  static_cast<Server*>(srv)->stats(submitted, completed, steps, frames);
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

int32_t mm_server_device_bytes(mm_server srv, int64_t* bytes) try {
  if (!srv || !bytes) return MM_ERR_INVALID_INPUT;
  *bytes = (int64_t)static_cast<Server*>(srv)->device_bytes();
  return MM_OK;
} catch (...) {   // no exception leaves the C ABI
  return MM_ERR_INTERNAL;
}

int32_t mm_server_destroy(mm_server srv) try {
  if (!srv) return MM_ERR_INVALID_INPUT;
  delete static_cast<Server*>(srv);
  return MM_OK;
} catch (const std::bad_alloc&) {   // no exception leaves the C ABI
  return MM_ERR_ALLOC_FAILED;
} catch (...) {
  return MM_ERR_INTERNAL;
}

}  // extern "C"
