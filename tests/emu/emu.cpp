// CPU emulator of the decoder's HIP kernels -- test infrastructure.
//
// Compiles the product's host planner (plan.cpp, parse.cpp) and the
// __host__ __device__ kernel phases (entropy_core.h, pixel_core.h) with g++ and
// runs every workgroup lane by lane, with the barriers between phases, so the
// kernels' indexing and the tile-link / lookback logic are checked against the
// oracle without a GPU.  `order` = 1 runs the decoupled lookback with every
// predecessor holding only its aggregate, which exercises the composition path.
#include <type_traits>
#include <stdint.h>
#include <string.h>

#include <memory>
#include <vector>

#include "entropy_core.h"
#include "pixel_core.h"
#include "plan.h"
#include "unstuff_core.h"
#include "prog_core.h"

using namespace cjd;

namespace {

struct Emu {
  Plan plan;
  std::vector<int16_t> coefs;
  std::vector<uint8_t> out;
  std::vector<uint32_t> status;  // device words, tagged with the launch epoch
  std::vector<uint32_t> bits;    // the host's status bits per image
  std::vector<uint32_t> ebuf;
  std::vector<BlockRec> recs;
  bool coefs_stale = false;  // records not yet expanded into coefs
  int fixes = 0;
  int fallbacks = 0;
};

// epoch of the emulated entropy launch
constexpr uint32_t kEpoch = 6;

struct EntRun {
  Plan* p;
  std::vector<LookbackRec> lb;
  std::vector<std::unique_ptr<EntShared>> shs;
  std::vector<EntCtx> ctx;
  std::vector<char> done;
  std::vector<char> failed;  // the item gave up in the lookback (no records)
  bool no_wait = false;      // a wait for an unpublished state gives up at once
  int fixes = 0;  // links decoded again in fix-up rounds (entries off the speculative chain)
};

// Link tasks: every queued task runs a round (task t on lane t), then the
// unfinished ones are packed, as in the kernel.
void run_links(EntShared& sh, const EntCtx& c) {
  for (;;) {
    const int n = sh.ntask;
    if (!n) return;
    std::vector<Task> T((size_t)n);
    for (int t = 0; t < n; ++t) T[(size_t)t] = task_get(sh, t);
    sh.nnext = 0;
    for (int t = 0; t < n; ++t) {
      Lane L;
      int j;
      if (!task_round(sh, t, c, T[(size_t)t], L, j)) task_put(sh, sh.nnext++, j, L);
    }
    sh.ntask = sh.nnext;
  }
}

void queue_links(EntShared& sh, const std::vector<char>& task) {
  sh.ntask = 0;
  for (int t = 0; t < kWG; ++t)
    if (task[(size_t)t]) task_new(sh, sh.ntask++, t);
}

// Fix-up rounds (each round: all lanes read, then the lanes that need it
// relink); returns the number of relinks.
int fix_rounds(EntShared& sh, const EntCtx& c) {
  int total = 0;
  for (;;) {
    std::vector<uint32_t> P(kWG), S(kWG);
    std::vector<char> need(kWG), task(kWG);
    int n = 0;
    for (int t = 0; t < kWG; ++t) n += (need[(size_t)t] = ent_fix_need(sh, t, c, P[(size_t)t], S[(size_t)t])) ? 1 : 0;
    if (!n) return total;
    total += n;
    for (int t = 0; t < kWG; ++t) task[(size_t)t] = need[(size_t)t] && link_begin(sh, c, t, P[(size_t)t], S[(size_t)t]);
    queue_links(sh, task);
    run_links(sh, c);
  }
}

void scan_counts(EntShared& sh, const EntCtx& c) {
  for (int t = 0; t < kWG; ++t) ent_scan_init(sh, t, c);
  for (int d = 1; d < c.J; d <<= 1) {  // each step: all lanes read, then all write
    std::vector<int32_t> v((size_t)kWG * 4);
    for (int t = 0; t < kWG; ++t) ent_scan_read(sh, t, d, &v[(size_t)t * 4]);
    for (int t = 0; t < c.J; ++t) ent_scan_write(sh, t, &v[(size_t)t * 4]);
  }
}

void count_all(EntShared& sh, const EntCtx& c) {
  for (int t = 0; t < kWG; ++t) ent_count(sh, t, c);
}

// Lookback, entry application and inclusive publication of item i, resolving
// predecessors first only when the kernel would have waited for them (or, with
// no_wait, giving up as a wait that ran out of time does).
void resolve(EntRun& R, size_t i) {
  if (R.done[i]) return;
  EntShared& sh = *R.shs[i];
  const EntCtx& c = R.ctx[i];
  if (c.lead) {
    auto wait = [&](uint32_t q, bool incl) -> int {
      auto ready = [&] {
        const int st = R.lb[q].status;
        return ((uint32_t)st >> 2) == kEpoch && (!incl || (st & 3) >= kLookIncl);
      };
      if (!ready()) {
        if (R.no_wait) return -1;
        resolve(R, q);
      }
      return R.lb[q].status;
    };
    uint32_t p = 0, s = 0;
    int32_t cc[4] = {0, 0, 0, 0};
    auto load = [&](uint32_t q) { return R.lb[q]; };
    if (!lookback_walk(R.p->items.data(), c.sg.first_item, (uint32_t)i, c.it.prev_item, wait, load, p, s, cc)) {
      R.lb[i].status = lookback_word(kEpoch, kLookFail);
      R.failed[i] = 1;
      R.done[i] = 1;
      return;
    }
    sh.entry_p = p;
    sh.entry_s = s;
    for (int k = 0; k < 4; ++k) sh.entry_c[k] = cc[k];
    if (ent_entry_differs(sh, c)) {
      sh.xp[0] = p;
      sh.xs[0] = (uint8_t)s;
      R.fixes += fix_rounds(sh, c);
      count_all(sh, c);
      scan_counts(sh, c);
    }
  }
  for (int t = 0; t < kWG; ++t) ent_absolute(sh, t, c);
  ent_inclusive(sh, c, R.lb[i]);
  R.lb[i].status = lookback_word(kEpoch, kLookIncl);
  R.done[i] = 1;
}

// k_expand, block by block (blocks the serial decoder wrote stay).
void expand(Emu& E) {
  const uint64_t ew = E.plan.entry_words;
  for (int64_t b = 0; b < E.plan.total_blocks; ++b) {
    const BlockRec r = E.recs[(size_t)b];
    if (r.nw & kRecDense) continue;
    int16_t* B = E.coefs.data() + b * 64;
    for (int k = 0; k < 64; ++k) B[k] = 0;
    const uint32_t n0 = r.nw & 255u, n = n0 < 64 ? n0 : 64, W = 16 * (1 + ((r.nw >> 8) & 3u));
    // a record whose groups would run past the entry buffer stages no AC entries
    const bool ok = n == 0 || entry_word(r.eidx, W, n - 1) - ((r.eidx + n - 1) & 3u) + 4 <= ew;
    for (uint32_t i = 0; ok && i < n; ++i) {
      const uint32_t e = E.ebuf[entry_word(r.eidx, W, i)];
      if (e != 0xFFFFFFFFu) B[(e >> 16) & 63] = (int16_t)(e & 0xFFFFu);
    }
    B[0] = r.dc;
  }
}

// Entry buffer left by the previous emulated launch (carry-over mode): a
// launch must not depend on what an earlier batch left in its regions.
std::vector<uint32_t> g_carry;
bool g_carry_on = false;
uint64_t g_lds_seed = 0;  // nonzero: every work group's LDS starts as random bytes
int g_ext = 0;  // CJD_EXT_* accepted by the planner (emu_ext)
int g_plant_ck = 0;       // nonzero: checkpoint words of every g_plant_ck-th tile claim too many DC entries

void run_entropy(Emu& E, int order) {
  Plan& p = E.plan;
  const size_t ni = p.items.size();
  E.ebuf.assign((size_t)p.entry_words, 0x7EADBEEFu);
  if (g_carry_on)
    for (size_t k = 0; k < E.ebuf.size() && k < g_carry.size(); ++k) E.ebuf[k] = g_carry[k];
  EntRun R;
  R.p = &p;
  // records left by an older launch (epoch kEpoch - 1, inclusive), which the
  // lookback must not take
  R.lb.resize(ni);
  memset(R.lb.data(), 0x5A, ni * sizeof(LookbackRec));
  for (auto& r : R.lb) r.status = lookback_word(kEpoch - 1, 2);
  R.shs.resize(ni);
  R.ctx.resize(ni);
  R.done.assign(ni, 0);
  R.failed.assign(ni, 0);
  R.no_wait = order == 2;
  // speculative pass, links, fix-up, counts, scan and aggregate publication
  // for every item
  for (size_t i = 0; i < ni; ++i) {
    R.shs[i].reset(new EntShared);
    EntShared& sh = *R.shs[i];
    // the work group's LDS as a previous occupant of the CU left it
    if (g_lds_seed == ~0ull) {
      // left uninitialised (a memory-sanitizer build reports any read of it)
    } else if (g_lds_seed) {
      uint64_t x = g_lds_seed * 0x9E3779B97F4A7C15ull + i;
      uint8_t* b = reinterpret_cast<uint8_t*>(&sh);
      for (size_t k = 0; k < sizeof(sh); ++k) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        b[k] = (uint8_t)x;
      }
    } else {
      memset(&sh, 0, sizeof(sh));
    }
    sh.lovf = 0;  // what the kernel's lane 0 sets before the first barrier
    sh.look_ok = 1;
    EntCtx& c = R.ctx[i];
    c.init(p.wptr, p.images.data(), p.segs.data(), p.tables.data(), E.ebuf.data(), p.items[i]);
    memcpy(sh.lut, &p.luts[(size_t)c.sg.tset * kLutEntries], sizeof(sh.lut));
    memcpy(sh.dz, kDeZigZag, 64);
    for (int k = 0; k < 4; ++k) canon_fill(sh.canon[k], c.tabs[k]);
    for (int k = 0; k < c.nh; ++k) {
      sh.tsel[2 * k] = (uint16_t)(seg_ac(c.sg, *c.im, k) * kL1Size);
      sh.tsel[2 * k + 1] = (uint16_t)(seg_dc(c.sg, *c.im, k) * kL1Size);
    }
    for (int t = 0; t < kWG; ++t) ent_spec(sh, t, c);
    if (g_plant_ck)  // failure injection: checkpoint words this pass did not write
      for (int t = 0; t < c.J; ++t)
        if ((c.tfirst + (uint32_t)t) % (uint32_t)g_plant_ck == (uint32_t)g_plant_ck / 2)
          for (int m = 0; m < sh.nck[t]; ++m) c.ckpts(t)[2 * m + 1] |= 0x3FFFu << 18;
    std::vector<char> task(kWG);
    for (int t = 0; t < kWG; ++t) task[(size_t)t] = ent_phase2(sh, t, c);
    queue_links(sh, task);
    run_links(sh, c);
    R.fixes += fix_rounds(sh, c);
    count_all(sh, c);
    scan_counts(sh, c);
    if (c.lead) {
      ent_aggregate(sh, c, R.lb[i]);
      R.lb[i].status = lookback_word(kEpoch, 1);
    }
  }
  // segment-first items resolve from the known start; then the rest in
  // order (predecessors inclusive) or backwards (predecessors aggregate-only).
  // order 2: backwards with no waiting and the segment-first items last, so
  // every item after a segment's first gives up (a failed lookback)
  // order >= 3: a random order (seeded by `order`), so that walks meet a mix
  // of inclusive and aggregate-only predecessors, as uneven progress on the
  // GPU gives them
  if (order >= 3) {
    std::vector<size_t> perm(ni);
    for (size_t k = 0; k < ni; ++k) perm[k] = k;
    uint64_t x = (uint64_t)order * 0x9E3779B97F4A7C15ull;
    for (size_t k = ni; k > 1; --k) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      std::swap(perm[k - 1], perm[(size_t)(x % k)]);
    }
    for (size_t k = 0; k < ni; ++k) resolve(R, perm[k]);
  } else {
    if (order != 2)
      for (size_t i = 0; i < ni; ++i)
        if (!R.ctx[i].lead) resolve(R, i);
    for (size_t k = 0; k < ni; ++k) resolve(R, order ? ni - 1 - k : k);
  }
  E.fixes = R.fixes;
  if (g_carry_on) g_carry = E.ebuf;
  for (size_t i = 0; i < ni; ++i) {
    if (R.failed[i]) {
      status_or(E.status.data() + R.ctx[i].im->error_index, kStatusUnresolved, kEpoch);
      continue;
    }
    for (int t = 0; t < kWG; ++t) ent_records(*R.shs[i], t, R.ctx[i], E.recs.data(), E.status.data(), kEpoch);
  }
  E.coefs_stale = true;
}

void run_pixel(Emu& E, int mode, size_t first, size_t count);

template <class SH>
void pixel_item(Emu& E, SH& sh, int mode, const PixCtx& c) {
  Plan& p = E.plan;
  memset(&sh, 0, sizeof(SH));
  for (int t = 0; t < kPixThreads; ++t) pix_stage_cos(sh, t);
  for (int t = 0; t < kPixThreads; ++t) pix_stage_zero(sh, t, c, p.qtabs.data());
  for (int t = 0; t < kPixThreads; ++t)
    pix_stage_blocks(sh, t, c, E.recs.data(), E.ebuf.data(), (uint64_t)p.entry_words, E.coefs.data());
  if constexpr (std::is_same<SH, PixSharedMfma>::value) {
    pix_idct_mfma_emul(sh, c);
  } else if (mode == kIdctInt) {
    for (int t = 0; t < kPixThreads; ++t) pix_rows<kIdctInt>(sh, t, c);
    for (int t = 0; t < kPixThreads; ++t) pix_cols<kIdctInt>(sh, t, c);
  } else {
    for (int t = 0; t < kPixThreads; ++t) pix_rows<kIdctFloat>(sh, t, c);
    for (int t = 0; t < kPixThreads; ++t) pix_cols<kIdctFloat>(sh, t, c);
  }
  for (int t = 0; t < kPixThreads; ++t) pix_planes(sh, t, c, E.out.data());
  for (int t = 0; t < kPixThreads; ++t) pix_colour(sh, t, c, E.out.data());
}

void run_pixel(Emu& E, int mode, size_t first, size_t count) {
  Plan& p = E.plan;
  std::unique_ptr<PixShared> sh(new PixShared);
  std::unique_ptr<PixSharedMfma> shm(new PixSharedMfma);
  for (size_t i = first; i < first + count; ++i) {
    PixCtx c;
    c.init(p.images.data(), p.pix[i]);
    if (mode == kIdctMfma) pixel_item(E, *shm, mode, c);
    else pixel_item(E, *sh, mode, c);
  }
}

// k_unstuff (unstuff.hip) lane by lane, a barrier between phases: every
// chunk of stuffed scan bytes into the word buffer (already 0xFF-filled).
void emu_k_unstuff(const UnstuffChunk* chunks, size_t nchunks, const uint8_t* raw, const SegmentDesc* segs,
                   size_t nsegs, uint8_t* out, size_t out_bytes, bool* oob) {
  std::unique_ptr<UnstuffShared> shp(new UnstuffShared);
  UnstuffShared& sh = *shp;
  std::vector<uint8_t> fn(kUnstuffLanes);
  std::vector<uint32_t> brk(kUnstuffLanes), tail(kUnstuffLanes);
  std::vector<char> consumed(kUnstuffLanes);
  for (size_t ci = 0; ci < nchunks; ++ci) {
    const UnstuffChunk& c = chunks[ci];
    memset(&sh, 0x5A, sizeof(sh));
    for (int l = 0; l < kUnstuffLanes; ++l) unstuff_load(sh, l, c, raw);
    unstuff_load_next(sh, c, raw);
    for (int l = 0; l < kUnstuffLanes; ++l) unstuff_carry(sh, l, c);
    for (int d = 1; d < kUnstuffLanes; d <<= 1) {
      for (int l = 0; l < kUnstuffLanes; ++l) fn[(size_t)l] = unstuff_fn_read(sh, l, d);
      for (int l = 0; l < kUnstuffLanes; ++l) sh.fn[l] = fn[(size_t)l];
    }
    for (int l = 0; l < kUnstuffLanes; ++l) consumed[(size_t)l] = unstuff_carry_in(sh, l, c);
    for (int l = 0; l < kUnstuffLanes; ++l) unstuff_count(sh, l, c, consumed[(size_t)l] != 0);
    for (int d = 1; d < kUnstuffLanes; d <<= 1) {
      for (int l = 0; l < kUnstuffLanes; ++l) unstuff_count_read(sh, l, d, brk[(size_t)l], tail[(size_t)l]);
      for (int l = 0; l < kUnstuffLanes; ++l) {
        sh.brk[l] = brk[(size_t)l];
        sh.tail[l] = tail[(size_t)l];
      }
    }
    // byte stores and segment reads, bounds-checked
    struct Out {
      uint8_t* p;
      size_t n;
      bool* oob;
      uint8_t dummy = 0;
      uint8_t& operator[](uint64_t k) {
        if (k < n) return p[k];
        *oob = true;
        return dummy;
      }
    } o{out, out_bytes, oob};
    auto base = [&](uint32_t sgi) -> uint64_t {
      if (sgi < nsegs) return segs[sgi].word_base * 4u;
      *oob = true;
      return 0;
    };
    if (unstuff_single_run(sh)) {
      const uint64_t dst0 = base(c.seg) + c.dst;
      const uint32_t lead = (uint32_t)(dst0 & 15u);
      for (int l = 0; l < kUnstuffLanes; ++l) unstuff_gather(sh, l, c, consumed[(size_t)l] != 0, lead);
      for (int l = 0; l < kUnstuffLanes; ++l) unstuff_flush(sh, l, sh.tail[kUnstuffLanes - 1], lead, dst0, o);
    } else {
      for (int l = 0; l < kUnstuffLanes; ++l) unstuff_write(sh, l, c, consumed[(size_t)l] != 0, o, base);
    }
  }
}

}  // namespace

extern "C" {

void* emu_run(int n, const uint8_t* const* files, const size_t* sizes, int idct_mode, int tile_bits,
              int want_planes, int order, int serial, int item_tiles, int ck_bits, int* errors) {
  Emu* E = new Emu;
  E->plan.tile_bits = tile_bits ? (uint32_t)tile_bits : 4096u;
  E->plan.ck_bits = (uint32_t)ck_bits;
  if (item_tiles) E->plan.item_tiles = (uint32_t)item_tiles;
  E->plan.want_planes = want_planes;
  E->plan.ext = g_ext;
  plan_batch(E->plan, n, files, sizes, errors);
  E->coefs.assign((size_t)E->plan.total_blocks * 64, (int16_t)0x5555);
  BlockRec junk{0, 0, 0x5555};  // eidx, nw, dc
  E->recs.assign((size_t)E->plan.total_blocks, junk);
  // as staging does: images with one-component scans start with zero records
  // (the blocks no scan codes keep them)
  for (const ImageRec& r : E->plan.recs)
    if (!r.error && r.f.num_scans > 1) {
      std::fill(E->recs.begin() + r.coef_base, E->recs.begin() + r.coef_base + r.f.num_blocks, BlockRec{0, 0, 0});
      std::fill(E->coefs.begin() + r.coef_base * 64, E->coefs.begin() + (r.coef_base + r.f.num_blocks) * 64,
                (int16_t)0);
    }
  E->out.assign(E->plan.out_bytes, 0xAB);
  // status words of an older launch with every bit set: this launch must
  // not read them as its own
  E->status.assign((size_t)n, (kEpoch - 1) << 8 | 0xFFu);
  // progressive images (k_prog_prepare, then k_prog level by level; within a
  // level in reverse task order, which must not matter)
  {
    const Plan& P = E->plan;
    for (int64_t i = 0; i < P.prog_blocks; ++i)
      prog_prepare_block(P.pimgs.data(), (uint32_t)P.pimgs.size(), i, E->coefs.data(), E->recs.data());
    for (size_t l = 0; l + 1 < P.plevel.size(); ++l)
      for (uint32_t t = P.plevel[l + 1]; t-- > P.plevel[l];)
      {
        const SegmentDesc& sg = P.segs[P.ptasks[t].seg];
        static ProgShared sh;  // the workgroup's LDS
        prog_task(sh, P.wptr, P.images.data(), sg, P.ptasks[t], P.tables.data() + (size_t)sg.tset * 4,
                  E->coefs.data(), E->status.data(), kEpoch);
      }
  }
  if (serial) {
    for (auto& s : E->plan.segs)
      ent_serial_segment(E->plan.wptr, E->plan.images.data(), s, E->plan.tables.data(), E->coefs.data(),
                         E->recs.data(), E->status.data(), kEpoch);
  } else {
    run_entropy(*E, order);
  }
  E->bits.assign((size_t)n, 0);
  for (int i = 0; i < n; ++i) {
    ImageRec& r = E->plan.recs[(size_t)i];
    if (r.error) continue;
    uint32_t& st = E->bits[(size_t)i];
    st = status_bits(E->status[(size_t)i], kEpoch);
    // as the engine: unresolved images and tiled syntax / truncation errors
    // are decoded again serially, and the serial verdict is reported
    if (!serial && !r.f.progressive &&
        (st & (kStatusUnresolved | kStatusInconsistent | kStatusSyntax | kStatusTruncated))) {
      // the serial redo is a launch of its own (next epoch)
      ++E->fallbacks;
      for (uint32_t s = 0; s < r.seg_count; ++s)
        ent_serial_segment(E->plan.wptr, E->plan.images.data(), E->plan.segs[r.seg_first + s],
                           E->plan.tables.data(), E->coefs.data(), E->recs.data(), E->status.data(), kEpoch + 1);
      st = status_bits(E->status[(size_t)i], kEpoch + 1) | kStatusFallback;
    }
    if (st & (kStatusSyntax | kStatusTruncated)) r.error = CJD_SYNTAX_ERROR;
    if (errors) errors[i] = r.error;
  }
  run_pixel(*E, idct_mode, 0, E->plan.pix.size());
  return E;
}

// what: 0 pixels, 1 coefficients, 2+c plane c; returns bytes copied or -1
long emu_get(void* h, int i, int what, void* dst, size_t cap) {
  Emu* E = static_cast<Emu*>(h);
  const ImageRec& r = E->plan.recs[(size_t)i];
  if (r.error) return -1;
  const uint8_t* src;
  size_t n;
  if (what == 0) {
    src = E->out.data() + r.out_offset;
    n = r.out_bytes;
  } else if (what == 1) {
    if (E->coefs_stale) {
      expand(*E);
      E->coefs_stale = false;
    }
    src = reinterpret_cast<const uint8_t*>(E->coefs.data() + r.coef_base * 64);
    n = (size_t)r.f.num_blocks * 128;
  } else {
    const int c = what - 2;
    if (c >= r.f.num_channels || r.plane_offset[c] == ~0ull) return -1;
    src = E->out.data() + r.plane_offset[c];
    n = (size_t)r.f.comp_stride[c] * r.f.comp_plane_height[c];
  }
  if (n > cap) return -1;
  memcpy(dst, src, n);
  return (long)n;
}

// [items, links decoded again in fix-up rounds, fallbacks, segments]
void emu_stats(void* h, long* out4) {
  Emu* E = static_cast<Emu*>(h);
  out4[0] = (long)E->plan.items.size();
  out4[1] = E->fixes;
  out4[2] = E->fallbacks;
  out4[3] = (long)E->plan.segs.size();
}

// per-image status bits of the last run (before the host maps them to codes)
void emu_status(void* h, uint32_t* out, int n) {
  Emu* E = static_cast<Emu*>(h);
  for (int i = 0; i < n && i < (int)E->bits.size(); ++i) out[i] = E->bits[(size_t)i];
}

void emu_free(void* h) { delete static_cast<Emu*>(h); }

// Device unstuffing of a batch against the host's: the batch planned both
// ways, the device plan's chunks run through the emulated k_unstuff into a
// 0xFF-filled word buffer.  Copies both word buffers out (cap words each);
// returns the word count, -1 when the plans disagree on it or cap is short,
// -2 when a byte store or segment read would leave its buffer.
long emu_unstuff_files(int n, const uint8_t* const* files, const size_t* sizes, uint32_t* out_host, uint32_t* out_dev,
                       size_t cap, int* errors) {
  Plan a, d;
  d.unstuff_device = 1;
  plan_batch(a, n, files, sizes, errors);
  plan_batch(d, n, files, sizes, nullptr);
  if (a.nwords != d.nwords || a.nwords > cap || d.wptr != nullptr) return -1;
  std::vector<uint32_t> w(d.nwords, 0xFFFFFFFFu);
  bool oob = false;
  emu_k_unstuff(d.chunks.data(), d.chunks.size(), d.rptr, d.segs.data(), d.segs.size(),
                reinterpret_cast<uint8_t*>(w.data()), w.size() * 4, &oob);
  if (oob) return -2;
  memcpy(out_host, a.wptr, a.nwords * 4);
  memcpy(out_dev, w.data(), w.size() * 4);
  return (long)a.nwords;
}

// The same on one raw scan (any bytes), with segments laid out as the
// planner lays them out: -3 on an out-of-sequence RSTn (both refuse it).
long emu_unstuff_bytes(const uint8_t* src, size_t len, int split, uint32_t* out_host, uint32_t* out_dev, size_t cap) {
  std::vector<size_t> bounds;
  std::vector<UnstuffChunk> chunks;
  if (!unstuff_segments(src, len, split != 0, &bounds, nullptr, nullptr, &chunks)) return -3;
  const size_t nseg = bounds.size() - 1;
  std::vector<SegmentDesc> segs(nseg);
  std::vector<size_t> dst_off(nseg);
  size_t nw = 0;
  for (size_t k = 0; k < nseg; ++k) {
    const size_t dlen = bounds[k + 1] - bounds[k];
    segs[k].word_base = nw;
    dst_off[k] = nw * 4;
    nw += (((dlen + 3) / 4 + kSegPadWords) + 3) & ~(size_t)3;
  }
  if (nw > cap) return -1;
  std::vector<uint32_t> h(nw, 0xFFFFFFFFu), w(nw, 0xFFFFFFFFu);
  unstuff_segments(src, len, split != 0, nullptr, reinterpret_cast<uint8_t*>(h.data()), dst_off.data(), nullptr);
  bool oob = false;
  emu_k_unstuff(chunks.data(), chunks.size(), src, segs.data(), segs.size(), reinterpret_cast<uint8_t*>(w.data()),
                w.size() * 4, &oob);
  if (oob) return -2;
  memcpy(out_host, h.data(), nw * 4);
  memcpy(out_dev, w.data(), nw * 4);
  return (long)nw;
}

// Failure injection: after each speculative pass the checkpoint words of every
// `every`-th tile claim more DC entries than the pass decoded (0 = off).
void emu_plant_ck(int every) { g_plant_ck = every; }

// Initial LDS of every emulated work group: random bytes from `seed` (0 =
// zeroed, ~0 = left uninitialised).
void emu_lds_seed(unsigned long long seed) { g_lds_seed = seed; }

// Extensions (CJD_EXT_*) the planner of the next runs accepts.
void emu_ext(int ext) { g_ext = ext; }

// Carry-over mode: every launch starts with the entry buffer the previous
// launch left (on = 1), instead of a fixed fill; 0 turns it off and forgets.
void emu_carry(int on) {
  g_carry_on = on != 0;
  if (!on) g_carry.clear();
}

}  // extern "C"
