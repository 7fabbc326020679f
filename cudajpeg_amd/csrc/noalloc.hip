// Allocation-free single-image decode (cjd_decode_noalloc): the path of
// cjd_batch_stage + cjd_batch_decode + cjd_batch_finish for one image, with
// every device buffer the decode needs -- unstuffed scan, tables,
// descriptors, work lists, the tiled decoder's entry regions, block records,
// dense coefficients for the serial redo, and the lookback scratch -- carved
// out of one file-scope device array instead of hipMalloc'd buffers.  For
// callers that cannot allocate device memory.  The kernels are the batch
// engine's.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "../../include/cudajpeg_amd.h"
#include "cjd_internal.h"
#include "launch.h"
#include "plan.h"

using namespace cjd;

namespace {

constexpr size_t kNoAllocBytes = (size_t)256 << 20;  // CJD_NOALLOC_SCRATCH_BYTES
constexpr int kMaxDevices = 16;

}  // namespace

// Device scratch of the allocation-free decode.  In host code this name is a
// host-side stand-in, not a device address: host code uses the per-device
// address from hipGetSymbolAddress (NoAllocDev::base) only.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

struct NoAllocDev {
  std::mutex m;       // one call at a time per device (a call ends synchronised)
  bool ready = false;
  uint8_t* base = nullptr;  // this device's address of `scratch`
  Plan plan;          // host planner state, reused from call to call
  uint32_t status[kScratchHead + 4];
};

NoAllocDev g_dev[kMaxDevices];
std::mutex g_init;

// Set-up on `device` (current): the device address of the scratch array.
int noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_init);
  NoAllocDev& d = g_dev[device];
  if (d.ready) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CJD_DEVICE_MEM_ERROR;
  d.base = static_cast<uint8_t*>(p);
  d.ready = true;
  return 0;
}

// Regions of the scratch for one planned image (byte offsets, 256-aligned).
struct Layout {
  size_t words, images, segs, items, pix, tables, luts, qtabs, btab, ebuf, recs, coefs, sc, end;
  size_t sc_words, status_words;
};

size_t take(size_t& at, size_t bytes) {
  const size_t o = at;
  at += (bytes + 255) & ~(size_t)255;
  return o;
}

Layout layout_of(const Plan& p) {
  Layout L{};
  size_t at = 0;
  L.words = take(at, p.nwords * 4);
  L.images = take(at, p.images.size() * sizeof(ImageDesc));
  L.segs = take(at, p.segs.size() * sizeof(SegmentDesc));
  L.items = take(at, p.items.size() * sizeof(WorkItem));
  L.pix = take(at, p.pix.size() * sizeof(PixItem));
  L.tables = take(at, p.tables.size() * sizeof(HuffTable));
  L.luts = take(at, p.luts.size() * 2);
  L.qtabs = take(at, p.qtabs.size() * 2);
  L.btab = take(at, p.images.size() * (size_t)kMfmaTabVecs * 16);
  L.ebuf = take(at, (size_t)p.entry_words * 4);
  L.recs = take(at, (size_t)p.total_blocks * sizeof(BlockRec));
  L.coefs = take(at, (size_t)p.total_blocks * 128);
  L.status_words = 4;  // one image, rounded to 4 words
  L.sc_words = kScratchHead + L.status_words + p.items.size() * (sizeof(LookbackRec) / 4);
  L.sc = take(at, L.sc_words * 4);
  L.end = at;
  return L;
}

template <class T>
hipError_t put(uint8_t* base, size_t off, const std::vector<T>& v, hipStream_t s) {
  if (v.empty()) return hipSuccess;
  return hipMemcpyAsync(base + off, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace

extern "C" int cjd_decode_noalloc(int device, const uint8_t* data, size_t size, int idct_mode, void* frame,
                                  size_t frame_cap, void* stream, size_t* scratch_needed) {
  if (device < 0 || device >= kMaxDevices || !data || idct_mode < 0 || idct_mode > 2) return CJD_PROGRAMMER_ERROR;
  if (hipSetDevice(device) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  if (int e = noalloc_setup(device)) return e;
  NoAllocDev& d = g_dev[device];
  std::lock_guard<std::mutex> lk(d.m);
  hipStream_t s = (hipStream_t)stream;
  Plan& p = d.plan;
  int err = 0;
  const uint8_t* files[1] = {data};
  const size_t sizes[1] = {size};
  plan_batch(p, 1, files, sizes, &err);
  if (err) return err;
  const Layout L = layout_of(p);
  if (scratch_needed) *scratch_needed = L.end;
  if (L.end > kNoAllocBytes) return CJD_OOM_ERROR;  // more scratch than the array holds
  const ImageRec& r = p.recs[0];
  if (r.out_offset != 0) return CJD_PROGRAMMER_ERROR;
  // the pixel kernel's 16-byte row chunks are aligned relative to the frame
  if (!frame || ((uintptr_t)frame & 15) || frame_cap < r.out_bytes) return CJD_PROGRAMMER_ERROR;
  uint8_t* const B = d.base;
  const auto* images = reinterpret_cast<const ImageDesc*>(B + L.images);
  const auto* segs = reinterpret_cast<const SegmentDesc*>(B + L.segs);
  const auto* items = reinterpret_cast<const WorkItem*>(B + L.items);
  const auto* pix = reinterpret_cast<const PixItem*>(B + L.pix);
  const auto* tables = reinterpret_cast<const HuffTable*>(B + L.tables);
  const auto* words = reinterpret_cast<const uint32_t*>(B + L.words);
  const auto* luts = reinterpret_cast<const uint16_t*>(B + L.luts);
  const auto* qtabs = reinterpret_cast<const uint16_t*>(B + L.qtabs);
  auto* btab = reinterpret_cast<uint4*>(B + L.btab);
  auto* ebuf = reinterpret_cast<uint32_t*>(B + L.ebuf);
  auto* recs = reinterpret_cast<BlockRec*>(B + L.recs);
  auto* coefs = reinterpret_cast<int16_t*>(B + L.coefs);
  auto* sc = reinterpret_cast<uint32_t*>(B + L.sc);
  uint32_t* status = sc + kScratchHead;
  auto* lb = reinterpret_cast<LookbackRec*>(status + L.status_words);
  auto* ticket = reinterpret_cast<unsigned long long*>(sc);
  auto* done = reinterpret_cast<unsigned long long*>(sc + 2 * kLaunchCounters);
  uint8_t* out = static_cast<uint8_t*>(frame);
  const uint32_t nitems = (uint32_t)p.items.size(), npix = (uint32_t)p.pix.size();
  // every call starts from a cleared head, status word and lookback records
  // (the previous call's contents stay in the array): ticket base 0, epoch 1
  hipError_t he = hipSuccess;
  auto ok = [&](hipError_t x) {
    if (he == hipSuccess) he = x;
  };
  if (p.nwords) ok(hipMemcpyAsync(B + L.words, p.wptr, p.nwords * 4, hipMemcpyHostToDevice, s));
  ok(put(B, L.images, p.images, s));
  ok(put(B, L.segs, p.segs, s));
  ok(put(B, L.items, p.items, s));
  ok(put(B, L.pix, p.pix, s));
  ok(put(B, L.tables, p.tables, s));
  ok(put(B, L.luts, p.luts, s));
  ok(put(B, L.qtabs, p.qtabs, s));
  if (he != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  ok(launch_clear(sc, L.sc_words, s));
  if (idct_mode == kIdctMfma) ok(launch_mfma_tables(images, 1, qtabs, btab, s));
  ok(launch_entropy(words, images, segs, items, 0, nitems, tables, luts, ebuf, recs, status, lb, ticket, done, 0, 1,
                    kDefaultWaitTicks, s));
  if (he != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  // status word and counters; the serial decoder redoes an unresolved image
  if (hipMemcpyAsync(d.status, sc, sizeof(d.status), hipMemcpyDeviceToHost, s) != hipSuccess) return CJD_DEVICE_MEM_ERROR;
  if (hipStreamSynchronize(s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  uint64_t head[kScratchHead / 2];
  memcpy(head, d.status, sizeof(head));
  uint32_t st = status_bits(d.status[kScratchHead], 1);
  if (head[0] != nitems || head[kLaunchCounters] != nitems) st |= kStatusUnresolved;
  // as cjd_batch_finish: an unresolved image, or a tiled syntax or
  // truncation error, is decoded again by the tiled decoder in a launch of
  // its own (epoch 2, the redo counter), then in ordered launches (epoch 3,
  // the next counter: no work item waits for another of its launch) and, if
  // that still reports anything, by the serial decoder (epoch 4), whose
  // verdict is reported
  const uint32_t redo_bits = kStatusUnresolved | kStatusInconsistent | kStatusSyntax | kStatusTruncated;
  auto read_status = [&](uint32_t epoch_of) -> int {
    if (hipMemcpyAsync(d.status + kScratchHead, status, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
      return CJD_DEVICE_MEM_ERROR;
    if (he != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
    st = status_bits(d.status[kScratchHead], epoch_of);
    return 0;
  };
  if (st & redo_bits) {
    ok(launch_entropy(words, images, segs, items, 0, nitems, tables, luts, ebuf, recs, status, lb,
                      ticket + kRedoCounter, done + kRedoCounter, 0, 2, kDefaultWaitTicks, s));
    if (int e = read_status(2)) return e;
  }
  if (st & redo_bits) {
    for (uint32_t a = 0, base = 0; a < nitems;) {
      const uint32_t e = ordered_launch_end(p.items, a, nitems);
      ok(launch_entropy(words, images, segs, items, a, e, tables, luts, ebuf, recs, status, lb,
                        ticket + kRedoCounter + 1, done + kRedoCounter + 1, base, 3, kDefaultWaitTicks, s));
      base += e - a;
      a = e;
    }
    if (int e = read_status(3)) return e;
  }
  if (st & redo_bits) {
    ok(launch_entropy_serial(words, images, segs, r.seg_count, tables, coefs, recs, status, 4, s));
    if (int e = read_status(4)) return e;
  }
  if (st & (kStatusSyntax | kStatusTruncated)) return CJD_SYNTAX_ERROR;
  ok(launch_pixel(idct_mode, images, pix, npix, recs, ebuf, (uint64_t)p.entry_words, coefs, qtabs, btab, out, s));
  if (he != hipSuccess) return CJD_KERNEL_LAUNCH_ERROR;
  return hipStreamSynchronize(s) == hipSuccess ? 0 : CJD_KERNEL_LAUNCH_ERROR;
}

extern "C" size_t cjd_noalloc_scratch_bytes(void) { return kNoAllocBytes; }
