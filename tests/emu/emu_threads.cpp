// Threaded CPU harness of the tiled entropy kernel -- test infrastructure.
//
// Runs the body of k_entropy (entropy.hip) for every work item with one host
// thread per lane (kWG threads) and a barrier wherever the kernel has one, so
// that lanes really run at the same time between barriers: built with
// ThreadSanitizer, any unsynchronised conflicting access between lanes (to
// the work group's shared state or to the entry buffer) is reported.  Items
// run one after another, so every lookback finds its predecessors' inclusive
// states.  Checks: coefficients (expanded from the block records) against a
// reference array given by the caller.
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include <memory>
#include <thread>
#include <vector>

#include "entropy_core.h"
#include "plan.h"

using namespace cjd;

namespace {

struct Group {
  pthread_barrier_t bar;
  Group() { pthread_barrier_init(&bar, nullptr, kWG); }
  ~Group() { pthread_barrier_destroy(&bar); }
  void sync() { pthread_barrier_wait(&bar); }
};

struct Launch {
  Plan* p;
  std::vector<uint32_t>* ebuf;
  std::vector<BlockRec>* recs;
  std::vector<uint32_t>* status;
  std::vector<LookbackRec>* lb;
  uint32_t epoch;
};

void run_links(Group& g, EntShared& sh, int tid, const EntCtx& c) {
  for (;;) {
    g.sync();
    const int n = sh.ntask;
    if (!n) break;
    Task T{};
    if (tid < n) T = task_get(sh, tid);
    if (tid == 0) sh.nnext = 0;
    g.sync();
    if (tid < n) {
      Lane L;
      int j;
      if (!task_round(sh, tid, c, T, L, j)) task_put(sh, __atomic_fetch_add(&sh.nnext, 1, __ATOMIC_RELAXED), j, L);
    }
    g.sync();
    if (tid == 0) sh.ntask = sh.nnext;
  }
}

void queue_links(Group& g, EntShared& sh, int tid, bool task, bool need = false) {
  if (tid == 0) {
    sh.ntask = 0;
    sh.nneed = 0;
  }
  g.sync();
  if (task) task_new(sh, __atomic_fetch_add(&sh.ntask, 1, __ATOMIC_RELAXED), tid);
  if (need) __atomic_fetch_add(&sh.nneed, 1, __ATOMIC_RELAXED);
}

void fix_rounds(Group& g, EntShared& sh, int tid, const EntCtx& c) {
  for (;;) {
    uint32_t p = 0, s = 0;
    const bool need = ent_fix_need(sh, tid, c, p, s);
    g.sync();
    const bool task = need && link_begin(sh, c, tid, p, s);
    queue_links(g, sh, tid, task, need);
    g.sync();
    const int n = sh.ntask, changed = sh.nneed;
    if (!n) {
      if (!changed) break;
      continue;
    }
    run_links(g, sh, tid, c);
  }
}

void scan_counts(Group& g, EntShared& sh, int tid, const EntCtx& c) {
  ent_scan_init(sh, tid, c);
  g.sync();
  for (int d = 1; d < c.J; d <<= 1) {
    int32_t v[4];
    ent_scan_read(sh, tid, d, v);
    g.sync();
    if (tid < c.J) ent_scan_write(sh, tid, v);
    g.sync();
  }
}

// One lane of one work group: the k_entropy body.
void lane_main(Group& g, EntShared& sh, const Launch& X, uint32_t item_id, int tid) {
  Plan& P = *X.p;
  EntCtx c;
  c.init(P.wptr, P.images.data(), P.segs.data(), P.tables.data(), X.ebuf->data(), P.items[item_id]);
  if (tid == 0) {
    sh.lovf = 0;
    sh.look_ok = 1;
    memcpy(sh.lut, &P.luts[(size_t)c.sg.tset * kLutEntries], sizeof(sh.lut));
    memcpy(sh.dz, kDeZigZag, 64);
  }
  if (tid < 4) canon_fill(sh.canon[tid], c.tabs[tid]);
  if (tid < c.nh) {
    sh.tsel[2 * tid] = (uint16_t)(seg_ac(c.sg, *c.im, tid) * kL1Size);
    sh.tsel[2 * tid + 1] = (uint16_t)(seg_dc(c.sg, *c.im, tid) * kL1Size);
  }
  g.sync();
  ent_spec(sh, tid, c);
  g.sync();
  queue_links(g, sh, tid, ent_phase2(sh, tid, c));
  run_links(g, sh, tid, c);
  fix_rounds(g, sh, tid, c);
  ent_count(sh, tid, c);
  g.sync();
  scan_counts(g, sh, tid, c);
  LookbackRec* lb = X.lb->data();
  LookbackRec* me = lb + item_id;
  if (tid == 0 && c.lead) {
    LookbackRec mine{};
    ent_aggregate(sh, c, mine);
    memcpy(reinterpret_cast<uint32_t*>(me) + kRecAgg0, reinterpret_cast<uint32_t*>(&mine) + kRecAgg0,
           4 * (kRecAgg1 - kRecAgg0));
    __atomic_store_n(&me->status, lookback_word(X.epoch, kLookAgg), __ATOMIC_RELEASE);
    auto wait = [&](uint32_t q, bool incl) -> int {
      const int st = __atomic_load_n(&lb[q].status, __ATOMIC_ACQUIRE);
      if (((uint32_t)st >> 2) == X.epoch && (!incl || (st & 3) >= kLookIncl)) return st;
      return -1;  // items run in order: a predecessor is always published
    };
    auto load = [&](uint32_t q) { return lb[q]; };
    uint32_t p = 0, s = 0;
    int32_t cc[4] = {0, 0, 0, 0};
    if (lookback_walk(P.items.data(), c.sg.first_item, item_id, c.it.prev_item, wait, load, p, s, cc)) {
      sh.entry_p = p;
      sh.entry_s = s;
      for (int q = 0; q < 4; ++q) sh.entry_c[q] = cc[q];
    } else {
      sh.look_ok = 0;
    }
  }
  g.sync();
  if (!sh.look_ok) return;
  if (ent_entry_differs(sh, c)) {
    if (tid == 0) {
      sh.xp[0] = sh.entry_p;
      sh.xs[0] = (uint8_t)sh.entry_s;
    }
    g.sync();
    fix_rounds(g, sh, tid, c);
    ent_count(sh, tid, c);
    g.sync();
    scan_counts(g, sh, tid, c);
  }
  ent_absolute(sh, tid, c);
  g.sync();
  if (tid == 0) {
    LookbackRec mine{};
    ent_inclusive(sh, c, mine);
    memcpy(reinterpret_cast<uint32_t*>(me) + kRecIncl0, reinterpret_cast<uint32_t*>(&mine) + kRecIncl0,
           4 * (kRecIncl1 - kRecIncl0));
    __atomic_store_n(&me->status, lookback_word(X.epoch, kLookIncl), __ATOMIC_RELEASE);
  }
  ent_records(sh, tid, c, X.recs->data(), X.status->data(), X.epoch);
}

}  // namespace

extern "C" {

// Decodes n files with the threaded harness (tile_bits, item_tiles as in the
// batch options); writes each image's coefficients (int16 [blocks][64]) to
// coefs[i] when not NULL and its status bits to bits[i].  Returns the number
// of work items.
int emt_run(int n, const uint8_t* const* files, const size_t* sizes, int tile_bits, int item_tiles,
            int16_t* const* coefs, uint32_t* bits) {
  Plan P;
  P.tile_bits = tile_bits ? (uint32_t)tile_bits : 4096u;
  if (item_tiles) P.item_tiles = (uint32_t)item_tiles;
  std::vector<int> errs((size_t)n);
  plan_batch(P, n, files, sizes, errs.data());
  std::vector<uint32_t> ebuf((size_t)P.entry_words, 0x7EADBEEFu);
  std::vector<BlockRec> recs((size_t)P.total_blocks, BlockRec{0, 0, 0x5555});
  std::vector<uint32_t> status((size_t)n, 0);
  std::vector<LookbackRec> lb(P.items.size());
  memset(lb.data(), 0, lb.size() * sizeof(LookbackRec));
  Launch X{&P, &ebuf, &recs, &status, &lb, 3};
  for (uint32_t it = 0; it < (uint32_t)P.items.size(); ++it) {
    Group g;
    std::unique_ptr<EntShared> sh(new EntShared);
    memset(sh.get(), 0xA5, sizeof(EntShared));
    std::vector<std::thread> th;
    th.reserve(kWG);
    for (int t = 0; t < kWG; ++t) th.emplace_back([&, t] { lane_main(g, *sh, X, it, t); });
    for (auto& x : th) x.join();
  }
  for (int i = 0; i < n; ++i) {
    const ImageRec& r = P.recs[(size_t)i];
    bits[i] = r.error ? 0xFFFFFFFFu : status_bits(status[(size_t)i], 3);
    if (r.error || !coefs || !coefs[i]) continue;
    for (int64_t b = 0; b < r.f.num_blocks; ++b) {
      const BlockRec R = recs[(size_t)(r.coef_base + b)];
      int16_t* B = coefs[i] + b * 64;
      memset(B, 0, 128);
      const uint32_t n0 = R.nw & 255u, m = n0 < 64 ? n0 : 64, W = 16 * (1 + ((R.nw >> 8) & 3u));
      for (uint32_t k = 0; k < m; ++k) {
        const uint64_t w = entry_word(R.eidx, W, k);
        if (w >= ebuf.size()) break;
        const uint32_t e = ebuf[w];
        if (e != 0xFFFFFFFFu) B[(e >> 16) & 63] = (int16_t)(e & 0xFFFFu);
      }
      B[0] = R.dc;
    }
  }
  return (int)P.items.size();
}

}  // extern "C"
