// Batch engine of the MI355X baseline-JPEG decoder: host staging (parse,
// unstuff, compact Huffman tables, work lists), device-resident buffers and
// kernel orchestration behind the C ABI of include/cudajpeg_amd.h.
//
// Host side of what the reference does in decodeScanGPU (src/decodeScanGPU.cu:
// 104-151: table reshuffle, unstuffing, scan upload) and in
// iDCT_resample_colourTransform (src/pixelTransformGPU.cu:621-677), with the
// grow-only buffer policy of ensureMemSize (src/format.cu:81-110).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/cudajpeg_amd.h"
#include "cjd_internal.h"
#include "launch.h"
#include "parse_internal.h"
#include "plan.h"

using namespace cjd;

#include "engine_internal.h"

namespace {

// Decode in parts (cjd_batch_set_overlap): the most runs, and the default.
constexpr int kMaxParts = 4;
constexpr int kDefaultParts = 2;
// Redo of an image the tiled decoder could not decode: kRelaunches launches
// of the tiled decoder over the image's work items alone, then kOrdered
// ordered redos (below), then the serial decoder, which decodes a segment on
// one lane (tens of seconds for an 8192x8192 scan without restart markers).
constexpr int kRelaunches = 1;
constexpr int kOrdered = 2;

// Grow-only device buffer (src/format.cu:81-110 policy).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap && p) return 0;
    // a buffer that has to grow again takes half as much again: hipFree waits
    // for the whole device, so a serving loop should seldom pay for a growth
    size_t n = std::max<size_t>(bytes, 256);
    if (p) n = std::max(n, cap + cap / 2);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, n) != hipSuccess) {
      p = nullptr;
      return CJD_DEVICE_MEM_ERROR;
    }
    cap = n;
    return 0;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Grow-only pinned host buffer: the staging area of a batch's uploads, so
// that each upload is one DMA on the caller's stream (from pageable memory
// the runtime copies through small staging buffers of its own, one after
// another, and the call returns only when they are done).
struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t bytes) {
    if (bytes <= cap && p) return true;
    // room to grow, and half as much again as before when growing again
    const size_t n = std::max<size_t>(std::max(bytes + bytes / 8, p ? cap + cap / 2 : 0), 4096);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    cap = n;
    return true;
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace

struct cjd_batch {
  cjd_context* ctx = nullptr;
  int idct_mode = kIdctInt;
  int entropy_mode = CJD_ENTROPY_TILED;
  Plan plan;

  DevBuf d_words, d_images, d_segs, d_items, d_pix, d_tables, d_luts, d_qtabs, d_coefs, d_out, d_scratch;
  DevBuf d_btab;  // MFMA table operands per image (k_mfma_tables at staging)
  DevBuf d_ptasks, d_pimgs;  // progressive images (CJD_EXT_PROGRESSIVE): k_prog tasks and coefficient ranges
  // Pinned staging: the planner writes the scan words into h_words, the
  // descriptors and tables are packed into h_meta, and every upload is a DMA
  // from them.  up_ev follows the last stage's uploads: the next stage waits
  // for it before writing into these buffers again.
  PinnedBuf h_words, h_meta, h_raw;
  // Device unstuffing (cjd_batch_set_unstuff): the stuffed scans and their
  // chunk states go up instead of the words, and k_unstuff writes the words.
  int unstuff = CJD_UNSTUFF_DEVICE;
  DevBuf d_raw, d_chunks;
  hipEvent_t up_ev = nullptr;
  bool up_pending = false;
  DevBuf d_ebuf, d_recs;  // tiled decoder: entry regions and block records
  // The pixel kernel reads the block records; the dense coefficient buffer
  // is filled from them (k_expand) only when a caller asks for it.
  bool coefs_stale = false;
  hipStream_t last_stream = nullptr;
  // scratch: a head of kScratchHead words (per launch j of a decode, a 64-bit
  // work-item ticket counter and a 64-bit counter of finished work items),
  // then status_words per-image status words, then the lookback records.
  // Nothing in it is cleared between decodes: the counters only grow (launch j
  // takes its nitems tickets from ticket_base[j] on) and the status words and
  // records are tagged with the launch's epoch.
  size_t scratch_words = 0, status_words = 0;
  uint64_t ticket_base[kLaunchCounters] = {};
  uint32_t epoch = 0;          // of the last entropy launch
  bool ticket_unknown = false;  // a launch failed: the counters are cleared before the next
  uint64_t wait_ticks = kDefaultWaitTicks;  // limit of one lookback wait
  std::vector<uint32_t> h_status;  // status words of the last decode (host), after the head
  int64_t fallbacks = 0;
  int64_t protocol_redos = 0;  // checks after a decode that found a counter off and redid the batch

  // Captured decode (cjd_batch_capture): valid while graph_signature(b)
  // equals the one taken at capture.  After a replay the head's counters of
  // launch j hold graph_counts[j] and the status words kGraphEpoch's bits.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> graph_sig;
  uint64_t graph_counts[kLaunchCounters] = {};
  bool replayed = false;

  // Decode in parts (cjd_batch_set_overlap): the batch's images are cut into
  // `parts` runs at work-item boundaries split_item[j] / pixel-item
  // boundaries split_pix[j] (j = 0 .. nsplit, whole images); the entropy
  // launches of the runs go one after another on the decode stream, and the
  // pixel kernel of every run but the last runs on `aux` as soon as its
  // entropy launch is done.  nsplit = 0: no split for this batch.
  int parts = kDefaultParts;
  int nsplit = 0;
  uint32_t split_item[kMaxParts + 1] = {}, split_pix[kMaxParts + 1] = {};
  hipStream_t aux = nullptr;
  hipEvent_t ex[kMaxParts] = {};  // entropy of run j done (j < nsplit - 1); the last: aux done

  // timed decodes: 0 start, 3 before and 4 after the entropy kernel(s), 1
  // after the entropy stage, 2 after the pixels
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool timed = false;

  ~cjd_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (up_pending) (void)hipEventSynchronize(up_ev);  // no DMA may still read the pinned buffers
    if (up_ev) (void)hipEventDestroy(up_ev);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : ex)
      if (e) (void)hipEventDestroy(e);
    if (aux) (void)hipStreamDestroy(aux);
  }
};

extern "C" int cjd_context_create(int device, cjd_context** out) {
  if (!out) return CJD_PROGRAMMER_ERROR;
  if (hipSetDevice(device) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  auto* c = new (std::nothrow) cjd_context;
  if (!c) return CJD_OOM_ERROR;
  c->device = device;
  *out = c;
  return 0;
}

extern "C" void cjd_context_destroy(cjd_context* ctx) { delete ctx; }

extern "C" int cjd_batch_create(cjd_context* ctx, cjd_batch** out) {
  if (!ctx || !out) return CJD_PROGRAMMER_ERROR;
  auto* b = new (std::nothrow) cjd_batch;
  if (!b) return CJD_OOM_ERROR;
  b->ctx = ctx;
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete b;
      return CJD_DEVICE_MEM_ERROR;
    }
  if (hipEventCreateWithFlags(&b->up_ev, hipEventDisableTiming) != hipSuccess) {
    delete b;
    return CJD_DEVICE_MEM_ERROR;
  }
  b->plan.word_store = [b](size_t n) -> uint32_t* {
    return b->h_words.ensure(n * 4) ? static_cast<uint32_t*>(b->h_words.p) : nullptr;
  };
  b->plan.raw_store = [b](size_t n) -> uint8_t* {
    return b->h_raw.ensure(n) ? static_cast<uint8_t*>(b->h_raw.p) : nullptr;
  };
  *out = b;
  return 0;
}

extern "C" void cjd_batch_destroy(cjd_batch* b) { delete b; }

extern "C" int cjd_batch_set_options(cjd_batch* b, int idct_mode, int entropy_mode, int tile_bits, int want_planes) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  if (idct_mode < 0 || idct_mode > 2 || entropy_mode < 0 || entropy_mode > 1) return CJD_PROGRAMMER_ERROR;
  if (tile_bits && (tile_bits < 512 || tile_bits > 8192 || (tile_bits & 31))) return CJD_PROGRAMMER_ERROR;
  b->idct_mode = idct_mode;
  b->entropy_mode = entropy_mode;
  if (tile_bits) b->plan.tile_bits = (uint32_t)tile_bits;
  b->plan.want_planes = want_planes ? 1 : 0;
  return 0;
}

extern "C" int cjd_batch_set_tuning(cjd_batch* b, int ck_bits, int item_tiles) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  if (ck_bits < 0 || (ck_bits && ck_bits < 64) || item_tiles < 0 || item_tiles == 1 || item_tiles > kWG)
    return CJD_PROGRAMMER_ERROR;
  b->plan.ck_bits = (uint32_t)ck_bits;
  b->plan.item_tiles = item_tiles ? (uint32_t)item_tiles : (uint32_t)kWG;
  return 0;
}

// Upload on the stream.  The host vectors belong to the batch and stay
// unchanged until its next stage, which comes after this decode finished.
extern "C" int cjd_batch_set_tune_table(cjd_batch* b, int n, const cjd_tune_entry* entries) {
  if (!b || n < 0 || (n && !entries)) return CJD_PROGRAMMER_ERROR;
  std::vector<TuneEntry> t;
  for (int i = 0; i < n; ++i) {
    const cjd_tune_entry& x = entries[i];
    if (x.tile_bits < 512 || x.tile_bits > 8192 || (x.tile_bits & 31) || x.item_tiles < 2 || x.item_tiles > kWG ||
        x.ck_bits < 0 || (x.ck_bits && x.ck_bits < 64) || x.max_bits_per_block < 0)
      return CJD_PROGRAMMER_ERROR;
    t.push_back(TuneEntry{x.max_scan_bytes, (uint32_t)x.tile_bits, (uint32_t)x.ck_bits, (uint32_t)x.item_tiles,
                          (uint32_t)x.max_bits_per_block});
  }
  // by size, then density (0 = any density last)
  auto dens = [](const TuneEntry& a) { return a.max_bpb ? a.max_bpb : ~0u; };
  std::sort(t.begin(), t.end(), [&](const TuneEntry& a, const TuneEntry& c) {
    return a.max_scan_bytes != c.max_scan_bytes ? a.max_scan_bytes < c.max_scan_bytes : dens(a) < dens(c);
  });
  b->plan.tune = t;
  return 0;
}

static int upload_bytes(DevBuf& d, const void* src, size_t bytes, hipStream_t s) {
  if (int e = d.ensure(bytes)) return e;
  if (bytes && hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  return 0;
}

// The plan's descriptor and table arrays, each uploaded from its copy in the
// pinned meta buffer (256-byte aligned), or from the vector itself when no
// pinned memory could be had.
struct MetaPacker {
  cjd_batch* b;
  size_t total = 0;
  template <class T>
  void count(const std::vector<T>& v) {
    total += (v.size() * sizeof(T) + 255) & ~(size_t)255;
  }
  size_t off = 0;
  template <class T>
  int put(DevBuf& d, const std::vector<T>& v, bool pinned, hipStream_t s) {
    const size_t bytes = v.size() * sizeof(T);
    const void* src = v.data();
    if (pinned && bytes) {
      uint8_t* dst = static_cast<uint8_t*>(b->h_meta.p) + off;
      memcpy(dst, v.data(), bytes);
      src = dst;
      off += (bytes + 255) & ~(size_t)255;
    }
    return upload_bytes(d, src, bytes, s);
  }
};

// Runs of the decode in parts: cut j (1 .. parts-1) at the first image
// boundary at or after (j+1)/(parts+1) of the work items, where both work
// lists are in image order; cuts that coincide or leave a run without pixel
// items are dropped (no split at all for small batches).  The first run is
// the largest: its pixel kernel then has the rest's entropy decode to run
// beside, and the last run's exposed pixel kernel is short (on the mixed
// batch, 2 parts cut at 2/3: 4.70 ms per decode; at 1/2: 4.90 ms).
static void plan_split(cjd_batch* b) {
  const Plan& p = b->plan;
  b->nsplit = 0;
  const size_t ni = p.items.size(), np = p.pix.size();
  if (b->parts < 2 || ni < 64 || np < 2) return;
  auto img = [&](size_t i) { return p.segs[p.items[i].segment].image; };
  for (size_t i = 1; i < ni; ++i)
    if (img(i) < img(i - 1)) return;
  for (size_t i = 1; i < np; ++i)
    if (p.pix[i].image < p.pix[i - 1].image) return;
  int n = 0;
  b->split_item[0] = b->split_pix[0] = 0;
  for (int j = 1; j < b->parts; ++j) {
    size_t m = ni * (size_t)(j + 1) / (size_t)(b->parts + 1);
    if (m == 0) continue;
    while (m < ni && img(m) == img(m - 1)) ++m;
    if (m >= ni || m <= b->split_item[n]) continue;
    const uint32_t d = img(m);
    size_t q = b->split_pix[n];
    while (q < np && p.pix[q].image < d) ++q;
    if (q <= b->split_pix[n] || q >= np) continue;
    ++n;
    b->split_item[n] = (uint32_t)m;
    b->split_pix[n] = (uint32_t)q;
  }
  if (n == 0) return;
  b->split_item[n + 1] = (uint32_t)ni;
  b->split_pix[n + 1] = (uint32_t)np;
  b->nsplit = n + 1;
}

extern "C" int cjd_batch_set_extensions(cjd_batch* b, int ext) {
  if (!b || (ext & ~(CJD_EXT_DQT16 | CJD_EXT_NONINTERLEAVED | CJD_EXT_PROGRESSIVE))) return CJD_PROGRAMMER_ERROR;
  b->plan.ext = ext;
  return 0;
}

extern "C" int cjd_batch_set_unstuff(cjd_batch* b, int mode) {
  if (!b || (mode != CJD_UNSTUFF_HOST && mode != CJD_UNSTUFF_DEVICE)) return CJD_PROGRAMMER_ERROR;
  b->unstuff = mode;
  return 0;
}

extern "C" int cjd_batch_set_overlap(cjd_batch* b, int parts) {
  if (!b || parts < 0 || parts > kMaxParts) return CJD_PROGRAMMER_ERROR;
  b->parts = parts == 0 ? 1 : parts;
  plan_split(b);
  return 0;
}

extern "C" int cjd_batch_stage(cjd_batch* b, int n, const uint8_t* const* files, const size_t* sizes, int* errors) {
  const int e = cjd_batch_stage_async(b, n, files, sizes, errors, nullptr);
  if (e) return e;
  return hipStreamSynchronize(nullptr) == hipSuccess ? 0 : CJD_DEVICE_MEM_ERROR;
}

extern "C" int cjd_batch_stage_async(cjd_batch* b, int n, const uint8_t* const* files, const size_t* sizes,
                                     int* errors, void* stream) {
  if (!b || n < 0 || (n && (!files || !sizes))) return CJD_PROGRAMMER_ERROR;
  hipStream_t s = (hipStream_t)stream;
  if (hipSetDevice(b->ctx->device) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  // the last stage's DMAs read the pinned buffers the planner writes next
  if (b->up_pending) {
    if (hipEventSynchronize(b->up_ev) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
    b->up_pending = false;
  }
  b->plan.unstuff_device = b->unstuff == CJD_UNSTUFF_DEVICE;
  plan_batch(b->plan, n, files, sizes, errors);
  const Plan& P = b->plan;
  MetaPacker mp{b};
  mp.count(P.chunks);
  mp.count(P.images);
  mp.count(P.segs);
  mp.count(P.items);
  mp.count(P.pix);
  mp.count(P.tables);
  mp.count(P.luts);
  mp.count(P.qtabs);
  mp.count(P.ptasks);
  mp.count(P.pimgs);
  const bool pinned = b->h_meta.ensure(mp.total);
  int e = 0;
  if (P.unstuff_device) {
    if (!e) e = b->d_words.ensure(P.nwords * 4);
    if (!e) e = upload_bytes(b->d_raw, P.rptr, P.nraw, s);
    if (!e) e = mp.put(b->d_chunks, P.chunks, pinned, s);
  } else {
    if (!e) e = upload_bytes(b->d_words, P.wptr, P.nwords * 4, s);
  }
  if (!e) e = mp.put(b->d_images, P.images, pinned, s);
  if (!e) e = mp.put(b->d_segs, P.segs, pinned, s);
  if (!e && P.unstuff_device &&
      launch_unstuff(b->d_chunks.as<UnstuffChunk>(), (uint32_t)P.chunks.size(), b->d_raw.as<uint8_t>(),
                     b->d_segs.as<SegmentDesc>(), b->d_words.as<uint32_t>(), P.nwords, s) != hipSuccess)
    e = CJD_KERNEL_LAUNCH_ERROR;
  if (!e) e = mp.put(b->d_items, P.items, pinned, s);
  if (!e) e = mp.put(b->d_pix, P.pix, pinned, s);
  if (!e) e = mp.put(b->d_tables, P.tables, pinned, s);
  if (!e) e = mp.put(b->d_luts, P.luts, pinned, s);
  if (!e) e = mp.put(b->d_ptasks, P.ptasks, pinned, s);
  if (!e) e = mp.put(b->d_pimgs, P.pimgs, pinned, s);
  if (!e && b->plan.entry_words >= ((int64_t)1 << 32)) e = CJD_OOM_ERROR;  // records hold 32-bit entry indexes
  if (!e) e = b->d_ebuf.ensure((size_t)b->plan.entry_words * 4);
  if (!e) e = b->d_recs.ensure((size_t)b->plan.total_blocks * sizeof(BlockRec));
  // one-component scans do not code the blocks MCUs pad a component's grid
  // with (outside the image), and decodes write only coded blocks' records:
  // those images' records start as zero blocks
  for (const ImageRec& r : P.recs)
    if (!e && !r.error && r.f.num_scans > 1 && !r.f.progressive &&
        launch_clear(reinterpret_cast<uint32_t*>(b->d_recs.as<BlockRec>() + r.coef_base),
                     (uint64_t)r.f.num_blocks * (sizeof(BlockRec) / 4), s) != hipSuccess)
      e = CJD_KERNEL_LAUNCH_ERROR;
  if (!e) e = mp.put(b->d_qtabs, P.qtabs, pinned, s);
  if (!e) {
    if (hipEventRecord(b->up_ev, s) != hipSuccess) e = CJD_DEVICE_MEM_ERROR;
    else b->up_pending = true;
  }
  if (!e) e = b->d_btab.ensure(b->plan.images.size() * (size_t)kMfmaTabVecs * 16);
  if (!e && launch_mfma_tables(b->d_images.as<ImageDesc>(), (uint32_t)b->plan.images.size(), b->d_qtabs.as<uint16_t>(),
                               b->d_btab.as<uint4>(), s) != hipSuccess)
    e = CJD_KERNEL_LAUNCH_ERROR;
  if (!e) e = b->d_coefs.ensure((size_t)b->plan.total_blocks * 128);
  // images with one-component scans: zero coefficient blocks as well (the
  // serial decoder writes the coded blocks only)
  for (const ImageRec& r : P.recs)
    if (!e && !r.error && r.f.num_scans > 1 && !r.f.progressive &&
        launch_clear(reinterpret_cast<uint32_t*>(b->d_coefs.as<int16_t>() + r.coef_base * 64),
                     (uint64_t)r.f.num_blocks * 32, s) != hipSuccess)
      e = CJD_KERNEL_LAUNCH_ERROR;
  if (!e) e = b->d_out.ensure(b->plan.out_bytes);
  // scratch (layout at cjd_batch): a new allocation is cleared whole and
  // starts the ticket counter again; otherwise the status words and records
  // are cleared (this batch's layout differs from the last one's) and the
  // counter and the epochs go on.  The clear is ordered on the staging stream
  // like the uploads.
  b->status_words = ((size_t)n + 3) & ~(size_t)3;
  b->scratch_words = kScratchHead + b->status_words + b->plan.items.size() * (sizeof(LookbackRec) / 4);
  if (!e) {
    const bool fresh = !b->d_scratch.p || b->d_scratch.cap < b->scratch_words * 4;
    e = b->d_scratch.ensure(b->scratch_words * 4);
    if (!e && fresh) {
      for (auto& t : b->ticket_base) t = 0;
      b->epoch = 0;
    }
    const size_t from = fresh ? 0 : kScratchHead;
    if (!e && launch_clear(b->d_scratch.as<uint32_t>() + from, b->scratch_words - from, s) != hipSuccess)
      e = CJD_KERNEL_LAUNCH_ERROR;
  }
  b->h_status.assign(kScratchHead + b->status_words, 0);
  plan_split(b);
  return e;
}

// Epoch of the next entropy launch on b's scratch.  After a failed launch the
// ticket counter may not match ticket_base: the whole scratch is cleared
// (ordered on s) and counting starts again.  After kEpochMax launches the
// status words and records are cleared and the epochs start again at 1.
static int next_epoch(cjd_batch* b, hipStream_t s, uint32_t* epoch) {
  if (b->ticket_unknown) {
    if (launch_clear(b->d_scratch.as<uint32_t>(), b->scratch_words, s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
    for (auto& t : b->ticket_base) t = 0;
    b->epoch = 0;
    b->ticket_unknown = false;
  } else if (b->epoch >= kEpochMax) {
    if (launch_clear(b->d_scratch.as<uint32_t>() + kScratchHead, b->scratch_words - kScratchHead, s) != hipSuccess)
      return CJD_KERNEL_LAUNCH_ERROR;
    b->epoch = 0;
  }
  *epoch = ++b->epoch;
  return 0;
}

static uint32_t* status_of(cjd_batch* b) { return b->d_scratch.as<uint32_t>() + kScratchHead; }

// Ticket and finished-item counters of launch j of a decode.
static unsigned long long* ticket_of(cjd_batch* b, int j) {
  return reinterpret_cast<unsigned long long*>(b->d_scratch.as<uint32_t>()) + j;
}
static unsigned long long* done_of(cjd_batch* b, int j) {
  return reinterpret_cast<unsigned long long*>(b->d_scratch.as<uint32_t>() + 2 * kLaunchCounters) + j;
}

// Epoch of a decode's entropy launches: the next one for a direct decode;
// for a captured one the fixed kGraphEpoch, after a kernel that clears what
// the decode's launches count and tag.
static int decode_epoch(cjd_batch* b, hipStream_t s, bool captured, uint32_t* epoch) {
  if (!captured) {
    b->replayed = false;  // a direct decode after a replay is checked as a direct one
    return next_epoch(b, s, epoch);
  }
  *epoch = kGraphEpoch;
  return launch_decode_begin(b->d_scratch.as<uint32_t>(), (uint32_t)b->status_words,
                             (uint32_t)b->plan.items.size(), s) == hipSuccess
             ? 0
             : CJD_KERNEL_LAUNCH_ERROR;
}

// Progressive images: their scans (k_prog_prepare, then k_prog level by
// level) with the decode's epoch, before its entropy launches.
static int enqueue_prog(cjd_batch* b, hipStream_t s, uint32_t epoch) {
  const Plan& P = b->plan;
  if (!P.prog_blocks) return 0;
  return launch_prog(b->d_pimgs.as<ProgImage>(), (uint32_t)P.pimgs.size(), P.prog_blocks, b->d_ptasks.as<ProgTask>(),
                     P.plevel.data(), (uint32_t)(P.plevel.size() - 1), b->d_words.as<uint32_t>(),
                     b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(), b->d_tables.as<HuffTable>(),
                     b->d_coefs.as<int16_t>(), b->d_recs.as<BlockRec>(), status_of(b), epoch, s) == hipSuccess
             ? 0
             : CJD_KERNEL_LAUNCH_ERROR;
}

static int enqueue_entropy(cjd_batch* b, hipStream_t s, bool timed, bool captured = false) {
  uint32_t* status = status_of(b);
  LookbackRec* lb = reinterpret_cast<LookbackRec*>(status + b->status_words);
  const uint32_t nitems = (uint32_t)b->plan.items.size();
  if (timed && hipEventRecord(b->ev[0], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  uint32_t epoch = 0;
  if (int e = decode_epoch(b, s, captured, &epoch)) return e;
  if (int e = enqueue_prog(b, s, epoch)) return e;
  if (timed && hipEventRecord(b->ev[3], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  hipError_t he;
  if (b->entropy_mode == CJD_ENTROPY_SERIAL)
    he = launch_entropy_serial(b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(),
                               (uint32_t)b->plan.segs.size(), b->d_tables.as<HuffTable>(), b->d_coefs.as<int16_t>(),
                               b->d_recs.as<BlockRec>(), status, epoch, s);
  else {
    he = launch_entropy(b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(),
                        b->d_items.as<WorkItem>(), 0, nitems, b->d_tables.as<HuffTable>(), b->d_luts.as<uint16_t>(),
                        b->d_ebuf.as<uint32_t>(), b->d_recs.as<BlockRec>(), status, lb, ticket_of(b, 0),
                        done_of(b, 0), captured ? 0 : b->ticket_base[0], epoch, b->wait_ticks, s);
    if (captured) {
      b->graph_counts[0] = nitems;
    } else {
      b->ticket_base[0] += nitems;
      b->ticket_unknown = he != hipSuccess;
    }
    if (he == hipSuccess && timed && hipEventRecord(b->ev[4], s) != hipSuccess) he = hipErrorUnknown;
    b->coefs_stale = true;
  }
  b->last_stream = s;
  if (he != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  if (timed && hipEventRecord(b->ev[1], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  return 0;
}

static hipError_t pixel_range(cjd_batch* b, uint32_t first, uint32_t count, hipStream_t s) {
  return launch_pixel(b->idct_mode, b->d_images.as<ImageDesc>(), b->d_pix.as<PixItem>() + first, count,
                      b->d_recs.as<BlockRec>(), b->d_ebuf.as<uint32_t>(), (uint64_t)b->plan.entry_words,
                      b->d_coefs.as<int16_t>(), b->d_qtabs.as<uint16_t>(), b->d_btab.as<uint4>(),
                      b->d_out.as<uint8_t>(), s);
}

static int enqueue_pixel(cjd_batch* b, hipStream_t s, bool timed) {
  if (pixel_range(b, 0, (uint32_t)b->plan.pix.size(), s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  if (timed && hipEventRecord(b->ev[2], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  return 0;
}

// Decode in parts on s: the runs' entropy launches one after another, the
// pixel kernel of every run but the last on `aux` after its entropy launch,
// the last run's pixel kernel on s; s waits for aux at the end, so the decode
// is complete when s is.  The entropy launches share the epoch (their
// images, work items and records are disjoint) and take consecutive tickets.
static int ensure_aux(cjd_batch* b) {
  if (b->aux) return 0;
  if (hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking) != hipSuccess) {
    b->aux = nullptr;
    return CJD_DEVICE_MEM_ERROR;
  }
  for (auto& e : b->ex)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  return 0;
}

static int enqueue_split(cjd_batch* b, hipStream_t s, bool timed, bool captured = false) {
  const int k = b->nsplit;
  if (int e = ensure_aux(b)) return e;
  uint32_t* status = status_of(b);
  LookbackRec* lb = reinterpret_cast<LookbackRec*>(status + b->status_words);
  if (timed && hipEventRecord(b->ev[0], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  uint32_t epoch = 0;
  if (int e = decode_epoch(b, s, captured, &epoch)) return e;
  if (int e = enqueue_prog(b, s, epoch)) return e;
  if (timed && hipEventRecord(b->ev[3], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  b->coefs_stale = true;
  b->last_stream = s;
  for (int j = 0; j < k; ++j) {
    const uint32_t i0 = b->split_item[j], i1 = b->split_item[j + 1];
    const hipError_t he =
        launch_entropy(b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(),
                       b->d_items.as<WorkItem>(), i0, i1, b->d_tables.as<HuffTable>(), b->d_luts.as<uint16_t>(),
                       b->d_ebuf.as<uint32_t>(), b->d_recs.as<BlockRec>(), status, lb, ticket_of(b, j),
                       done_of(b, j), captured ? 0 : b->ticket_base[j], epoch, b->wait_ticks, s);
    if (captured) {
      b->graph_counts[j] = i1 - i0;
    } else {
      b->ticket_base[j] += i1 - i0;
    }
    if (he != hipSuccess) {
      if (!captured) b->ticket_unknown = true;
      return CJD_KERNEL_LAUNCH_ERROR;
    }
    if (j + 1 < k) {
      const uint32_t q0 = b->split_pix[j], q1 = b->split_pix[j + 1];
      if (hipEventRecord(b->ex[j], s) != hipSuccess || hipStreamWaitEvent(b->aux, b->ex[j], 0) != hipSuccess ||
          pixel_range(b, q0, q1 - q0, b->aux) != hipSuccess)
        return CJD_KERNEL_LAUNCH_ERROR;
    }
  }
  if (timed && (hipEventRecord(b->ev[4], s) != hipSuccess || hipEventRecord(b->ev[1], s) != hipSuccess))
    return CJD_KERNEL_LAUNCH_ERROR;
  const uint32_t ql = b->split_pix[k - 1], qe = b->split_pix[k];
  if (hipEventRecord(b->ex[kMaxParts - 1], b->aux) != hipSuccess || pixel_range(b, ql, qe - ql, s) != hipSuccess ||
      hipStreamWaitEvent(s, b->ex[kMaxParts - 1], 0) != hipSuccess)
    return CJD_KERNEL_LAUNCH_ERROR;
  if (timed && hipEventRecord(b->ev[2], s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  return 0;
}

static int enqueue(cjd_batch* b, hipStream_t s, bool timed, bool captured = false) {
  if (b->nsplit > 1 && b->entropy_mode == CJD_ENTROPY_TILED) return enqueue_split(b, s, timed, captured);
  int e = enqueue_entropy(b, s, timed, captured);
  return e ? e : enqueue_pixel(b, s, timed);
}

// Everything a captured decode bakes into its launches: buffers, sizes,
// runs, modes and the wait limit.  A replay is valid while this is unchanged.
static std::vector<uint64_t> graph_signature(const cjd_batch* b) {
  std::vector<uint64_t> v = {
      (uint64_t)b->d_words.p,  (uint64_t)b->d_images.p, (uint64_t)b->d_segs.p,    (uint64_t)b->d_items.p,
      (uint64_t)b->d_pix.p,    (uint64_t)b->d_tables.p, (uint64_t)b->d_luts.p,    (uint64_t)b->d_qtabs.p,
      (uint64_t)b->d_coefs.p,  (uint64_t)b->d_out.p,    (uint64_t)b->d_scratch.p, (uint64_t)b->d_btab.p,
      (uint64_t)b->d_ebuf.p,   (uint64_t)b->d_recs.p,   (uint64_t)b->aux,         b->status_words,
      b->plan.items.size(),    b->plan.pix.size(),      b->plan.segs.size(),      (uint64_t)b->plan.entry_words,
      (uint64_t)b->idct_mode,  (uint64_t)b->entropy_mode, b->wait_ticks,          (uint64_t)b->nsplit,
      (uint64_t)b->d_ptasks.p, (uint64_t)b->d_pimgs.p,  b->plan.pimgs.size(),     (uint64_t)b->plan.prog_blocks};
  for (uint32_t l : b->plan.plevel) v.push_back(l);
  for (int j = 0; j <= kMaxParts; ++j) {
    v.push_back(b->split_item[j]);
    v.push_back(b->split_pix[j]);
  }
  return v;
}

// Wait for the entropy stage; decode again with the serial decoder every image
// whose tile links did not resolve; map status bits to reference error codes.
// `redo` receives the images that were decoded again.
// Ordered redo of image r with the tiled decoder: its work items in launches
// [a, e) chosen so that no item waits in the lookback for an item of its own
// launch (every item of the launch either starts its segment or follows an
// item of an earlier launch), all with one epoch, one after another on s.  A
// lookback then finds its predecessor's inclusive state at once: there is no
// wait between work groups, at the cost of one launch per item of a segment's
// run (about 100 launches of a few microseconds for an 8192x8192 scan).
static int relaunch_ordered(cjd_batch* b, const ImageRec& r, uint32_t epoch, hipStream_t s) {
  uint32_t* status = status_of(b);
  LookbackRec* lb = reinterpret_cast<LookbackRec*>(status + b->status_words);
  const uint32_t end = r.item_first + r.item_count;
  for (uint32_t a = r.item_first; a < end;) {
    const uint32_t e = ordered_launch_end(b->plan.items, a, end);
    const hipError_t he = launch_entropy(
        b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(),
        b->d_items.as<WorkItem>(), a, e, b->d_tables.as<HuffTable>(), b->d_luts.as<uint16_t>(),
        b->d_ebuf.as<uint32_t>(), b->d_recs.as<BlockRec>(), status, lb, ticket_of(b, kRedoCounter),
        done_of(b, kRedoCounter), b->ticket_base[kRedoCounter], epoch, b->wait_ticks, s);
    b->ticket_base[kRedoCounter] += e - a;
    if (he != hipSuccess) {
      b->ticket_unknown = true;
      return CJD_KERNEL_LAUNCH_ERROR;
    }
    a = e;
  }
  return 0;
}

static int check_entropy(cjd_batch* b, hipStream_t s, std::vector<size_t>* redo) {
  // the counters of the head and the status words, in one copy
  if (hipMemcpyAsync(b->h_status.data(), b->d_scratch.p, b->h_status.size() * 4, hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return CJD_DEVICE_MEM_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  // After the decodes, launch j's ticket counter and finished-item counter
  // both equal the tickets handed out to it: every work group took one ticket
  // and every work item finished once.  Anything else means some item was not
  // decoded (its records may be stale): every image is decoded again by the
  // serial decoder, and the counters start again at the next decode.
  bool redo_all = false;
  const bool replayed = b->replayed;
  b->replayed = false;
  if (b->entropy_mode == CJD_ENTROPY_TILED) {
    uint64_t head[kScratchHead / 2];
    memcpy(head, b->h_status.data(), sizeof(head));
    for (int j = 0; j < kLaunchCounters; ++j) {
      const uint64_t want = replayed ? b->graph_counts[j] : b->ticket_base[j];
      if (head[j] != want || head[kLaunchCounters + j] != want) redo_all = true;
    }
    if (redo_all) ++b->protocol_redos;
  }
  // after a replay the counters no longer follow ticket_base, nor the words
  // the epochs: the next direct launch (a redo below included) clears them
  if (redo_all || replayed) b->ticket_unknown = true;
  const uint32_t epoch = replayed ? kGraphEpoch : b->epoch;
  uint32_t* words = b->h_status.data() + kScratchHead;
  for (size_t i = 0; i < b->status_words; ++i) words[i] = status_bits(words[i], epoch);
  // Every image the tiled decoder could not resolve, or for which it reports
  // a syntax error or a truncated scan, is decoded again: first by the tiled
  // decoder in a launch of its own over the image's work items alone, then in
  // ordered launches (relaunch_ordered), then, if those still report
  // anything, by the serial decoder (the reference's order), whose verdict is
  // the one reported -- a tiled error is never reported unless the serial
  // decoder agrees.  Each redo has an epoch of its own, so the image's status
  // word starts without bits.
  const bool tiled = b->entropy_mode == CJD_ENTROPY_TILED;
  const uint32_t redo_bits =
      tiled ? (kStatusUnresolved | kStatusInconsistent | kStatusSyntax | kStatusTruncated) : 0u;
  uint32_t* status = status_of(b);
  auto read_status = [&](size_t i, uint32_t epoch_of, uint32_t& bits) -> int {
    uint32_t w = 0;
    if (hipMemcpyAsync(&w, status + i, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
    if (hipStreamSynchronize(s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
    bits = status_bits(w, epoch_of);
    return 0;
  };
  for (size_t i = 0; i < b->plan.recs.size(); ++i) {
    ImageRec& r = b->plan.recs[i];
    r.decode_error = 0;
    if (r.error) continue;
    if (r.f.progressive) {  // k_prog's verdict is final (no redo)
      if (words[i] & (kStatusSyntax | kStatusTruncated)) r.decode_error = CJD_SYNTAX_ERROR;
      continue;
    }
    const uint32_t st = words[i] | (redo_all ? kStatusUnresolved : 0u);
    if (st & redo_bits) {
      const uint32_t first = (st & 0xFFu) << kStatusFirstShift;
      bool redone = false;
      uint32_t e2 = 0, bits = 0;
      for (int attempt = 0; attempt < kRelaunches && r.item_count && !redone; ++attempt) {
        if (int e = next_epoch(b, s, &e2)) return e;
        LookbackRec* lb = reinterpret_cast<LookbackRec*>(status + b->status_words);
        const hipError_t he = launch_entropy(
            b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(), b->d_segs.as<SegmentDesc>(),
            b->d_items.as<WorkItem>(), r.item_first, r.item_first + r.item_count, b->d_tables.as<HuffTable>(),
            b->d_luts.as<uint16_t>(), b->d_ebuf.as<uint32_t>(), b->d_recs.as<BlockRec>(), status, lb,
            ticket_of(b, kRedoCounter), done_of(b, kRedoCounter), b->ticket_base[kRedoCounter], e2, b->wait_ticks, s);
        b->ticket_base[kRedoCounter] += r.item_count;
        if (he != hipSuccess) {
          b->ticket_unknown = true;
          return CJD_KERNEL_LAUNCH_ERROR;
        }
        if (int e = read_status(i, e2, bits)) return e;
        redone = !(bits & redo_bits);
        if (redone) words[i] = bits | kStatusRelaunch | first;
      }
      for (int attempt = 0; attempt < kOrdered && r.item_count && !redone; ++attempt) {
        if (int e = next_epoch(b, s, &e2)) return e;
        if (int e = relaunch_ordered(b, r, e2, s)) return e;
        if (int e = read_status(i, e2, bits)) return e;
        redone = !(bits & redo_bits);
        if (redone) words[i] = bits | kStatusOrdered | first;
      }
      if (!redone) {
        if (int e = next_epoch(b, s, &e2)) return e;
        if (launch_entropy_serial(b->d_words.as<uint32_t>(), b->d_images.as<ImageDesc>(),
                                  b->d_segs.as<SegmentDesc>() + r.seg_first, r.seg_count, b->d_tables.as<HuffTable>(),
                                  b->d_coefs.as<int16_t>(), b->d_recs.as<BlockRec>(), status, e2, s) != hipSuccess)
          return CJD_KERNEL_LAUNCH_ERROR;
        if (int e = read_status(i, e2, bits)) return e;
        words[i] = bits | kStatusFallback | first;
      }
      b->coefs_stale = true;
      ++b->fallbacks;
      if (redo) redo->push_back(i);
    }
    if (words[i] & (kStatusSyntax | kStatusTruncated)) r.decode_error = CJD_SYNTAX_ERROR;
  }
  return 0;
}

extern "C" int cjd_batch_decode_scan(cjd_batch* b, void* stream) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  b->timed = false;
  int e = enqueue_entropy(b, (hipStream_t)stream, false);
  if (!e) e = check_entropy(b, (hipStream_t)stream, nullptr);
  return e;
}

extern "C" int cjd_batch_transform(cjd_batch* b, void* stream) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  return enqueue_pixel(b, (hipStream_t)stream, false);
}

extern "C" int cjd_batch_decode(cjd_batch* b, void* stream) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  b->timed = true;
  return enqueue(b, (hipStream_t)stream, true);
}

static void drop_graph(cjd_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_sig.clear();
}

// The decode as one HIP graph: k_decode_begin, the entropy launches with the
// fixed kGraphEpoch and ticket base 0, and the pixel launches (the decode in
// parts with its second stream joined in by events).  Nothing is run.
extern "C" int cjd_batch_capture(cjd_batch* b, void* stream) {
  if (!b || !stream || b->plan.recs.empty() || !b->d_scratch.p) return CJD_PROGRAMMER_ERROR;
  hipStream_t s = (hipStream_t)stream;
  drop_graph(b);
  for (auto& g : b->graph_counts) g = 0;  // kinds the captured decode does not launch expect no tickets
  if (b->nsplit > 1)
    if (int e = ensure_aux(b)) return e;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  const int e = enqueue(b, s, false, true);
  hipGraph_t g = nullptr;
  const hipError_t he = hipStreamEndCapture(s, &g);
  if (e || he != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return e ? e : CJD_KERNEL_LAUNCH_ERROR;
  }
  if (hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->gexec = nullptr;
    (void)hipGraphDestroy(g);
    return CJD_KERNEL_LAUNCH_ERROR;
  }
  b->graph = g;
  b->graph_sig = graph_signature(b);
  return 0;
}

// One replay of the captured decode on `stream`; cjd_batch_finish then checks
// it like a direct decode.  UNSUPPORTED_ERROR when the batch was re-staged or
// re-configured so that the graph's launches no longer match it.
extern "C" int cjd_batch_replay(cjd_batch* b, void* stream) {
  if (!b || !b->gexec) return CJD_PROGRAMMER_ERROR;
  if (graph_signature(b) != b->graph_sig) return CJD_UNSUPPORTED_ERROR;
  hipStream_t s = (hipStream_t)stream;
  if (hipGraphLaunch(b->gexec, s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  b->replayed = true;
  b->ticket_unknown = true;  // the replay reset the counters: a direct launch after it clears first
  b->timed = false;
  b->coefs_stale = true;
  b->last_stream = s;
  return 0;
}

extern "C" int cjd_batch_finish(cjd_batch* b, void* stream, int* errors) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  hipStream_t s = (hipStream_t)stream;
  std::vector<size_t> redo;
  int e = check_entropy(b, s, &redo);
  if (e) return e;
  for (size_t i : redo) {
    const ImageRec& r = b->plan.recs[i];
    if (pixel_range(b, r.pix_first, r.pix_count, s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  }
  if (!redo.empty() && hipStreamSynchronize(s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  if (errors)
    for (size_t i = 0; i < b->plan.recs.size(); ++i) errors[i] = b->plan.recs[i].code();
  return 0;
}

extern "C" int cjd_batch_fallbacks(cjd_batch* b) { return b ? (int)b->fallbacks : -1; }

extern "C" int cjd_batch_set_wait_limit(cjd_batch* b, int64_t microseconds) {
  if (!b || microseconds < 0 || microseconds > (int64_t)1 << 40) return CJD_PROGRAMMER_ERROR;
  b->wait_ticks = (uint64_t)microseconds * 100u;  // 100 MHz real-time counter
  return 0;
}

extern "C" int cjd_batch_protocol_redos(cjd_batch* b) { return b ? (int)b->protocol_redos : -1; }

extern "C" int cjd_batch_status(cjd_batch* b, int i, uint32_t* bits) {
  if (!b || !bits || i < 0 || (size_t)i >= b->plan.recs.size() || (size_t)i >= b->status_words)
    return CJD_PROGRAMMER_ERROR;
  *bits = b->h_status[kScratchHead + (size_t)i];
  return 0;
}

extern "C" int cjd_batch_timings(cjd_batch* b, float* ms_entropy, float* ms_pixel, float* ms_total) {
  if (!b || !b->timed) return CJD_PROGRAMMER_ERROR;
  if (hipEventSynchronize(b->ev[2]) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  float a = 0, c = 0, t = 0;
  (void)hipEventElapsedTime(&a, b->ev[0], b->ev[1]);
  (void)hipEventElapsedTime(&c, b->ev[1], b->ev[2]);
  (void)hipEventElapsedTime(&t, b->ev[0], b->ev[2]);
  if (ms_entropy) *ms_entropy = a;
  if (ms_pixel) *ms_pixel = c;
  if (ms_total) *ms_total = t;
  return 0;
}

extern "C" int cjd_batch_kernel_timings(cjd_batch* b, float* ms_entropy_kernel, float* ms_expand) {
  if (!b || !b->timed || b->entropy_mode != CJD_ENTROPY_TILED) return CJD_PROGRAMMER_ERROR;
  if (hipEventSynchronize(b->ev[1]) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  float a = 0, x = 0;
  (void)hipEventElapsedTime(&a, b->ev[3], b->ev[4]);
  (void)hipEventElapsedTime(&x, b->ev[4], b->ev[1]);
  if (ms_entropy_kernel) *ms_entropy_kernel = a;
  if (ms_expand) *ms_expand = x;
  return 0;
}

extern "C" int cjd_batch_image_info(cjd_batch* b, int i, cjd_frame_info* info) {
  if (!b || i < 0 || (size_t)i >= b->plan.recs.size() || !info) return CJD_PROGRAMMER_ERROR;
  *info = b->plan.recs[(size_t)i].f;
  return b->plan.recs[(size_t)i].code();
}

extern "C" int cjd_batch_device_ptr(cjd_batch* b, int i, int what, void** ptr, size_t* bytes) {
  if (!b || i < 0 || (size_t)i >= b->plan.recs.size()) return CJD_PROGRAMMER_ERROR;
  const ImageRec& r = b->plan.recs[(size_t)i];
  if (r.code()) return r.code();
  size_t n = 0;
  void* p = nullptr;
  if (what == CJD_OUT_PIXELS) {
    p = b->d_out.as<uint8_t>() + r.out_offset;
    n = r.out_bytes;
  } else if (what == CJD_OUT_COEFFICIENTS) {
    if (b->coefs_stale) {
      // blocks the serial decoder wrote are in place already; k_expand fills
      // the others from their records
      if (launch_expand(b->d_recs.as<BlockRec>(), b->d_ebuf.as<uint32_t>(), (uint64_t)b->plan.entry_words,
                        b->d_coefs.as<int16_t>(), (uint64_t)b->plan.total_blocks, b->last_stream) != hipSuccess ||
          hipStreamSynchronize(b->last_stream) != hipSuccess)
        return CJD_KERNEL_LAUNCH_ERROR;
      b->coefs_stale = false;
    }
    p = b->d_coefs.as<int16_t>() + r.coef_base * 64;
    n = (size_t)r.f.num_blocks * 128;
  } else if (what >= CJD_OUT_PLANE0 && what < CJD_OUT_PLANE0 + r.f.num_channels) {
    const int c = what - CJD_OUT_PLANE0;
    if (r.plane_offset[c] == ~0ull) return CJD_PROGRAMMER_ERROR;
    p = b->d_out.as<uint8_t>() + r.plane_offset[c];
    n = (size_t)r.f.comp_stride[c] * r.f.comp_plane_height[c];
  } else {
    return CJD_PROGRAMMER_ERROR;
  }
  if (ptr) *ptr = p;
  if (bytes) *bytes = n;
  return 0;
}

extern "C" int cjd_batch_download(cjd_batch* b, int i, int what, void* host, size_t cap) {
  void* p = nullptr;
  size_t n = 0;
  int e = cjd_batch_device_ptr(b, i, what, &p, &n);
  if (e) return e;
  if (n > cap) return CJD_PROGRAMMER_ERROR;
  return hipMemcpy(host, p, n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : CJD_DEVICE_MEM_ERROR;
}

extern "C" int cjd_batch_totals(cjd_batch* b, int64_t* images, int64_t* pixels, int64_t* scan_bytes, int64_t* blocks,
                                int64_t* out_bytes) {
  if (!b) return CJD_PROGRAMMER_ERROR;
  if (images) *images = (int64_t)b->plan.images.size();
  if (pixels) *pixels = b->plan.total_pixels;
  if (scan_bytes) *scan_bytes = b->plan.scan_bytes;
  if (blocks) *blocks = b->plan.total_blocks;
  if (out_bytes) {
    int64_t ob = 0;
    for (auto& r : b->plan.recs)
      if (!r.error) ob += (int64_t)r.out_bytes;
    *out_bytes = ob;
  }
  return 0;
}
