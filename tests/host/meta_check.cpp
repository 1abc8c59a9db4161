// Host-side checks of the fused kernel's work description (no GPU needed):
//  * per-tile prefetch records: a frame is marked for the plain interior load
//    only if it exists, lies wholly inside its clip and starts at an even
//    offset, and its packed-signal address is the clip's own frame start;
//  * the mel-block-to-wave assignment: every block exactly once, no wave over
//    its slot count, waves holding blocks form a prefix, fields decode.
// Built against libmetalmom.so by tests/test_host_logic.py; prints "ok".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../metalmom_amd/csrc/mm_engine.h"

using namespace mm;

static int fails = 0;
#define CHECK(c, ...)                        \
  do {                                       \
    if (!(c)) {                              \
      if (++fails < 20) {                    \
        std::fprintf(stderr, "FAIL %s: ", #c); \
        std::fprintf(stderr, __VA_ARGS__);   \
        std::fprintf(stderr, "\n");          \
      }                                      \
    }                                        \
  } while (0)

static void check_meta(const SpecRequest& r, const std::vector<int64_t>& lens, int tf) {
  BatchMeta m;
  build_meta(r, lens.data(), (int)lens.size(), tf, &m);
  const int64_t pad = r.center ? r.n_fft / 2 : 0;
  CHECK(m.tile_pf.size() == m.tile_clip.size(), "record count");
  int64_t in = 0, tiles = 0;
  for (size_t c = 0; c < lens.size(); ++c) {
    CHECK(m.clip_off[c] == in, "clip_off %zu", c);
    const int64_t T = frame_count(lens[c], r.n_fft, r.hop, r.center);
    CHECK(m.clip_frames[c] == T, "frames of clip %zu", c);
    tiles += (T + tf - 1) / tf;
    in += lens[c];
  }
  CHECK((int64_t)m.tile_clip.size() == tiles, "tile count");
  for (size_t t = 0; t < m.tile_clip.size(); ++t) {
    const int c = m.tile_clip[t];
    const int64_t f0 = m.tile_f0[t], T = m.clip_frames[c], len = m.clip_len[c], off = m.clip_off[c];
    const TilePf& p = m.tile_pf[t];
    CHECK((p.mask >> tf) == 0u, "mask bits past the tile (tile %zu)", t);
    for (int w = 0; w < tf; ++w) {
      const int64_t fr = f0 + w, s0 = fr * r.hop - pad;
      const bool want = fr < T && s0 >= 0 && s0 + r.n_fft <= len && (s0 & 1) == 0;
      const bool got = (p.mask >> w) & 1u;
      CHECK(want == got, "mask tile %zu frame %lld (n_fft %d hop %d len %lld)", t, (long long)fr, r.n_fft, r.hop,
            (long long)len);
      if (got) {
        const int64_t a = p.base + (int64_t)w * r.hop;
        CHECK(a == off + s0, "address tile %zu frame %lld: %lld vs %lld", t, (long long)fr, (long long)a,
              (long long)(off + s0));
        CHECK(a >= off && a + r.n_fft <= off + len, "range tile %zu frame %lld", t, (long long)fr);
      }
    }
  }
}

static void check_assign(const std::vector<int2>& band, int nw) {
  std::vector<int> out(kMaxMelBlocks, 7);
  const int nmh = (int)band.size();
  mel_block_assign(band.data(), nmh, nw, out.data());
  const int slots = kMaxMelBlocks / nw;
  std::vector<int> seen(nmh, 0), first(nmh, 0);
  for (int b = 1; b < nmh; ++b) first[b] = first[b - 1] + band[b - 1].y;
  int last_busy = -1;
  for (int w = 0; w < nw; ++w) {
    bool prefix = true;
    int nb = 0;
    for (int i = 0; i < slots; ++i) {
      const int v = out[w * slots + i];
      if (!slot_used(v)) {
        CHECK(v == 0, "empty slot %d of wave %d is not 0", i, w);
        prefix = false;
        continue;
      }
      CHECK(prefix, "slots of wave %d are not a prefix", w);
      ++nb;
      const int b = slot_block(v);
      CHECK(b < nmh, "block id %d", b);
      if (b < nmh) {
        seen[b] += 1;
        CHECK(slot_first_bin(v) == band[b].x && slot_chunks(v) == band[b].y, "decode of block %d", b);
        CHECK(slot_cache_first(v) == (first[b] & 255), "cache offset of block %d", b);
      }
    }
    if (nb > 0) {
      CHECK(last_busy == w - 1, "waves with blocks are not a prefix (wave %d)", w);
      last_busy = w;
    }
  }
  if (nmh <= kMaxMelBlocks) {
    for (int b = 0; b < nmh; ++b) CHECK(seen[b] == 1, "block %d placed %d times", b, seen[b]);
    CHECK(last_busy + 1 == (nmh < nw ? nmh : nw), "busy waves %d for %d blocks", last_busy + 1, nmh);
  } else {
    CHECK(last_busy == -1, "oversized configuration must leave every slot empty");
  }
}

// LDS filterbank cache sources: chunk order, offsets, and every 8-row source
// chunk inside the table [rows][fbs]
static void check_cache(const std::vector<int2>& band, int fbs, int ch) {
  const int nmh = (int)band.size();
  std::vector<int> src(kMaxCacheChunks + 1, -7);
  const int n = cache_sources(band.data(), nmh, fbs, ch, src.data());
  int total = 0;
  for (int b = 0; b < nmh; ++b) total += band[b].y;
  if (total > kMaxCacheChunks) {
    CHECK(n == 0, "oversized cache must be refused (%d chunks)", total);
    return;
  }
  CHECK(n == total, "cache chunks %d vs %d", n, total);
  const int rows = 16 * ((nmh * kMelBlock + 15) / 16);
  int i = 0;
  for (int b = 0; b < nmh; ++b)
    for (int c = 0; c < band[b].y; ++c, ++i) {
      CHECK(src[i] == kMelBlock * b * fbs + band[b].x + ch * c, "source of chunk %d", i);
      CHECK(band[b].x + ch * (c + 1) <= fbs, "chunk %d of block %d runs past the row", c, b);
      CHECK(kMelBlock * b + kMelBlock <= rows, "block %d rows", b);
    }
  CHECK(src[n] == -7, "write past the chunk list");
}

int main() {
  std::mt19937_64 rng(7);
  // framing: n_fft, hop, centre, tile width, odd and tiny clip lengths
  const int nffts[] = {256, 512, 1024, 2048, 4096};
  const int hops[] = {1, 160, 333, 512, 1024, 3000};
  for (int nf : nffts)
    for (int hop : hops)
      for (int center = 0; center < 2; ++center)
        for (int tf : {8, 16}) {
          SpecRequest r;
          r.n_fft = nf;
          r.win = nf;
          r.hop = hop;
          r.center = center;
          std::vector<int64_t> lens;
          for (int i = 0; i < 24; ++i) {
            const int64_t kind = rng() % 4;
            int64_t n = kind == 0 ? (int64_t)(rng() % 64) + 1
                      : kind == 1 ? nf - 1 + (int64_t)(rng() % 3)
                                  : (int64_t)(rng() % 200000) + nf;
            if (!center && n < nf) n = nf;   // uncentred framing needs a whole frame
            lens.push_back(n);
          }
          check_meta(r, lens, tf);
        }
  // block assignment: monotone (mel-like), random, constant and empty bands
  for (int nw : {8, 16})
    for (int nmh = 0; nmh <= kMaxMelBlocks + 2; ++nmh)
      for (int pattern = 0; pattern < 4; ++pattern) {
        std::vector<int2> band(nmh);
        for (int b = 0; b < nmh; ++b) {
          int n = pattern == 0 ? 1 + b * b / 8 : pattern == 1 ? (int)(rng() % 40) : pattern == 2 ? 3 : (b % 3 ? 0 : 5);
          band[b] = make_int2((int)(rng() % 2000) & ~7, n);
        }
        check_assign(band, nw);
        if (nw == 16) {
          // filterbank rows are fbs wide: a band of n chunks of ch bins fits
          const int fbs = 2048 + 64;
          for (int ch : {16, 32}) {
            std::vector<int2> bb = band;
            for (auto& x : bb) {   // bands as the table generator makes them
              x.x = (x.x & ~7) % 2048;
              x.y = std::min(x.y, (fbs - x.x) / ch);
            }
            check_cache(bb, fbs, ch);
          }
        }
      }
  if (fails) {
    std::fprintf(stderr, "%d failures\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
