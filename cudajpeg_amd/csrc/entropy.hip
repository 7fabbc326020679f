// Entropy-decode kernels of the MI355X baseline-JPEG decoder (gfx950).
//
// k_entropy runs the phases of entropy_core.h for one work item (a run of
// tiles of one restart segment) per 256-lane workgroup, with a decoupled
// lookback across workgroups, and leaves a record per block; k_expand turns
// the records into coefficient blocks; k_entropy_serial is the one-lane-per-
// segment decoder (CJD_ENTROPY_SERIAL, and images the tiled decoder hands
// back).  Together they replace
// decodeScanGPU (src/decodeScanGPU.cu:123-491) and the speculative kernels of
// src/entropyRLEdecodeGPU.cu:7-238.
#include <hip/hip_runtime.h>

#include "entropy_core.h"
#include "launch.h"

namespace cjd {

__device__ __forceinline__ int ld_status(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Lookback records are read and written word by word with agent-scope
// atomics: a record is written by one work group and read by others, on other
// XCDs, whose L2 caches are separate; an agent-scope acquire does not drop
// every line an L2 holds, so a plain load after it could return a copy of the
// line older than the writer's release (seen as a wrong entry state: syntax
// and truncation errors on valid scans, now and then, under concurrent
// decodes).  Atomic accesses at agent scope are coherent across the XCDs.
__device__ __forceinline__ LookbackRec ld_record(const LookbackRec* r) {
  LookbackRec v{};
  const uint32_t* src = reinterpret_cast<const uint32_t*>(r);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
  for (int k = kRecAgg0; k < kRecIncl1; ++k)
    dst[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return v;
}
__device__ __forceinline__ void st_record_words(LookbackRec* r, const LookbackRec& v, int w0, int w1) {
  uint32_t* dst = reinterpret_cast<uint32_t*>(r);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
  for (int k = 0; k < kRecWords; ++k)
    if (k >= w0 && k < w1) __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First kernel of a captured decode: zero the launch counters, the status
// words and the status word of every lookback record, so that the decode's
// entropy launches can take ticket base 0 and the fixed kGraphEpoch on every
// replay (a record that shows no state of this replay is waited for).
__global__ void k_decode_begin(uint32_t* __restrict__ scratch, uint32_t nstatus, uint32_t nrec) {
  const uint32_t words = kScratchHead + nstatus;
  const uint32_t n = words + nrec;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    if (k < words) scratch[k] = 0u;
    else scratch[words + (uint64_t)(k - words) * kRecWords] = 0u;
  }
}

__global__ void k_clear_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Wave64 primitives.  A wave's lanes claim slots of an LDS list with one
// atomic per wave: the ballot of the lanes that want one gives the wave's
// count (added by the lowest such lane) and each lane's rank among them
// (mbcnt).  Called by every lane of the wave (the ballot must see them all).
__device__ __forceinline__ int wave_claim(bool want, int32_t* counter) {
  const uint64_t m = __ballot(want);
  if (!m) return 0;  // wave-uniform
  const uint32_t below =
      __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const int leader = __builtin_ctzll(m);
  int base = 0;
  if ((int)(threadIdx.x & 63u) == leader) base = atomicAdd(counter, __builtin_popcountll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  return base + (int)below;
}

__device__ __forceinline__ void wave_count(bool pred, int32_t* counter) {
  const uint64_t m = __ballot(pred);
  if (m && (int)(threadIdx.x & 63u) == __builtin_ctzll(m)) atomicAdd(counter, __builtin_popcountll(m));
}

// Inclusive prefix sum over the 64 lanes of a wave in DPP steps: shifts of
// 1, 2, 4 and 8 lanes within each row of 16 (lanes shifted in read 0), then
// lane 15 of each row added into the next row (rows 1 and 3) and lane 31
// into rows 2 and 3.  Each step is one v_add_u32 with a DPP source.
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
  return x;
}

// Run the sh.ntask queued link tasks in rounds; after each round the
// unfinished ones are packed into the lowest lanes (all lanes call this).
__device__ __forceinline__ void run_links(EntShared& sh, int tid, const EntCtx& c) {
  for (;;) {
    __syncthreads();  // tasks and their count written
    const int n = sh.ntask;
    if (!n) break;
    Task T{};
    if (tid < n) T = task_get(sh, tid);
    if (tid == 0) sh.nnext = 0;
    __syncthreads();  // every task read before any is put back
    Lane L;
    int j = 0;
    const bool keep = tid < n && !task_round(sh, tid, c, T, L, j);
    const int idx = wave_claim(keep, &sh.nnext);
    if (keep) task_put(sh, idx, j, L);
    __syncthreads();
    if (tid == 0) sh.ntask = sh.nnext;
  }
}

// Queue a link task for tile tid when `task`; sh.nneed counts the lanes with
// `need` (all lanes call this; both counts are read after the next barrier).
__device__ __forceinline__ void queue_links(EntShared& sh, int tid, bool task, bool need = false) {
  if (tid == 0) {
    sh.ntask = 0;
    sh.nneed = 0;
  }
  __syncthreads();
  const int idx = wave_claim(task, &sh.ntask);
  if (task) task_new(sh, idx, tid);
  wave_count(need, &sh.nneed);
}

// Fix-up rounds (all lanes): relink every tile whose entry changed, until no
// tile changes.  The "did any tile change" decision is a count in LDS taken
// beside the task count, not a library work-group reduction.
__device__ __forceinline__ void fix_rounds(EntShared& sh, int tid, const EntCtx& c) {
  for (;;) {
    uint32_t p = 0, s = 0;
    const bool need = ent_fix_need(sh, tid, c, p, s);
    __syncthreads();  // reads of the previous state done
    const bool task = need && link_begin(sh, c, tid, p, s);
    queue_links(sh, tid, task, need);
    __syncthreads();
    const int n = sh.ntask, changed = sh.nneed;
    if (!n) {
      // a relink may have joined at once without a task: the exits moved, check again
      if (!changed) break;
      continue;
    }
    run_links(sh, tid, c);
  }
}

// Inclusive prefix sum of the per-tile counts into tc (all lanes): a DPP scan
// within each wave, then each lane adds the totals of the waves before its
// own -- one barrier, where a scan over the workgroup in LDS takes two per
// doubling step.  Each lane reads only its own tile's counts (written by it
// in ent_count); the barrier at the end publishes tc.
__device__ __forceinline__ void scan_counts(EntShared& sh, int tid, const EntCtx& c) {
  const int w = tid >> 6;
  int32_t v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = wave_incl_scan(tid < c.J ? sh.u.c.cnt[tid][q] : 0);
  if ((tid & 63) == 63)
    for (int q = 0; q < 4; ++q) sh.wsum[w][q] = v[q];
  __syncthreads();
  for (int k = 0; k < w; ++k)
    for (int q = 0; q < 4; ++q) v[q] += sh.wsum[k][q];
  if (tid < c.J)
    for (int q = 0; q < 4; ++q) sh.u.c.tc[tid][q] = v[q];
  __syncthreads();
}

__global__ void __launch_bounds__(kWG) k_entropy(const uint32_t* __restrict__ words,
                                                 const ImageDesc* __restrict__ images,
                                                 const SegmentDesc* __restrict__ segs,
                                                 const WorkItem* __restrict__ items, uint32_t item0,
                                                 uint32_t item_end, const HuffTable* __restrict__ tables,
                                                 const uint16_t* __restrict__ luts, uint32_t* __restrict__ ebuf,
                                                 BlockRec* __restrict__ recs, uint32_t* __restrict__ status,
                                                 LookbackRec* __restrict__ lb,
                                                 unsigned long long* __restrict__ ticket,
                                                 unsigned long long* __restrict__ done, uint64_t ticket_base,
                                                 uint32_t epoch, uint64_t wait_ticks) {
  __shared__ EntShared sh;
  const int tid = threadIdx.x;
  // Work items are taken in ticket order, so every predecessor an item waits
  // for in the lookback has already started.  The ticket counter is never
  // cleared and belongs to this launch of the decode alone: a launch over
  // items [item0, item_end) takes that many tickets from ticket_base on.
  if (tid == 0) {
    sh.item = item0 + (uint32_t)(atomicAdd(ticket, 1ull) - ticket_base);
    sh.lovf = 0;
    sh.look_ok = 1;
  }
  __syncthreads();
  const uint32_t item_id = sh.item;
  if (item_id < item0 || item_id >= item_end) return;
  EntCtx c;
  c.init(words, images, segs, tables, ebuf, items[item_id]);
  {
    const uint4* src = reinterpret_cast<const uint4*>(luts + (size_t)c.sg.image * kLutEntries);
    uint4* dst = reinterpret_cast<uint4*>(sh.lut);
    for (int i = tid; i < kLutEntries / 8; i += kWG) dst[i] = src[i];
    if (tid < 64) sh.dz[tid] = kDeZigZag[tid];
    if (tid < 4) canon_fill(sh.canon[tid], c.tabs[tid]);
    if (tid < c.nh) {
      sh.tsel[2 * tid] = (uint16_t)(c.im->slot_ac[tid] * kL1Size);
      sh.tsel[2 * tid + 1] = (uint16_t)(c.im->slot_dc[tid] * kL1Size);
    }
  }
  __syncthreads();
  ent_spec(sh, tid, c);
  __syncthreads();
  queue_links(sh, tid, ent_phase2(sh, tid, c));
  run_links(sh, tid, c);
  fix_rounds(sh, tid, c);
  ent_count(sh, tid, c);
  __syncthreads();
  scan_counts(sh, tid, c);
  LookbackRec* me = lb + item_id;
  LookbackRec mine;
  if (tid == 0 && c.lead) {
    ent_aggregate(sh, c, mine);
    st_record_words(me, mine, kRecAgg0, kRecAgg1);
    st_status(&me->status, lookback_word(epoch, kLookAgg));
    // decoupled lookback: compose predecessors' aggregates until one has its
    // inclusive state.  A record of an older launch does not match this
    // epoch; every wait ends, after wait_ticks of the real-time counter at
    // most, and then the item gives up.
    auto wait = [&](uint32_t q, bool incl) -> int {
      const int32_t* w = &lb[q].status;
      uint64_t t0 = 0;
      for (bool first = true;; first = false) {
        const int st = ld_status(w);
        if (((uint32_t)st >> 2) == epoch && (!incl || (st & 3) >= kLookIncl)) return st;
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (first) t0 = now;
        else if (now - t0 >= wait_ticks) return -1;
        __builtin_amdgcn_s_sleep(2);
      }
    };
    uint32_t p = 0, s = 0;
    int32_t cc[4] = {0, 0, 0, 0};
    auto load = [&](uint32_t q) { return ld_record(lb + q); };
    if (lookback_walk(items, c.sg.first_item, item_id, c.it.prev_item, wait, load, p, s, cc)) {
      sh.entry_p = p;
      sh.entry_s = s;
      for (int q = 0; q < 4; ++q) sh.entry_c[q] = cc[q];
    } else {
      // successors give up at this record; the host decodes the image again
      sh.look_ok = 0;
      status_or(status + c.im->error_index, kStatusUnresolved, epoch);
      st_status(&me->status, lookback_word(epoch, kLookFail));
    }
  }
  __syncthreads();
  if (!sh.look_ok) {
    if (tid == 0) atomicAdd(done, 1ull);
    return;
  }
  if (ent_entry_differs(sh, c)) {  // uniform: the chain is redone from the true entry
    if (tid == 0) {
      sh.xp[0] = sh.entry_p;
      sh.xs[0] = (uint8_t)sh.entry_s;
    }
    __syncthreads();
    fix_rounds(sh, tid, c);
    ent_count(sh, tid, c);
    __syncthreads();
    scan_counts(sh, tid, c);
  }
  ent_absolute(sh, tid, c);
  __syncthreads();
  if (tid == 0) {
    ent_inclusive(sh, c, mine);
    st_record_words(me, mine, kRecIncl0, kRecIncl1);
    st_status(&me->status, lookback_word(epoch, kLookIncl));
    atomicAdd(done, 1ull);  // read by the host after the launch
  }
  ent_records(sh, tid, c, recs, status, epoch);
}

// Block records -> quantised coefficients (natural order, absolute DC), 256
// consecutive blocks per workgroup, written as one contiguous 32 KB run
// (blocks the serial decoder wrote are left as they are; the pixel kernel
// reads records, so this runs when coefficients are asked for).  A block's AC
// entries are consecutive entries of its tile's stream, 16-byte groups of 4
// one stream row apart: a lane issues its block's group loads four at a
// time before using any.  The blocks are assembled in LDS at a stride of 33
// words, so that lanes writing the same coefficient index of their blocks hit
// different banks.
constexpr int kExpandStride = 33;  // LDS words per block (32 + 1 of padding)

__global__ void __launch_bounds__(256) k_expand(const BlockRec* __restrict__ recs, const uint32_t* __restrict__ ebuf,
                                                uint64_t ewords, int16_t* __restrict__ coefs, uint64_t nblocks) {
  __shared__ uint32_t blk[256 * kExpandStride];
  const int tid = threadIdx.x;
  const uint64_t b0 = (uint64_t)blockIdx.x * 256;
  uint32_t* mine = blk + tid * kExpandStride;
#pragma unroll
  for (int i = 0; i < 32; ++i) mine[i] = 0;
  const uint64_t b = b0 + tid;
  if (b < nblocks && !(recs[b].nw & kRecDense)) {
    const BlockRec r = recs[b];
    const uint32_t n0 = r.nw & 255u, n = n0 < 64 ? n0 : 64, W = 16 * (1 + ((r.nw >> 8) & 3u));
    const uint32_t ph = r.eidx & 3u, ng = n ? (ph + n + 3) >> 2 : 0u;
    const uint64_t g0 = r.eidx & ~3ull, row = 4ull * W;
    int16_t* B = reinterpret_cast<int16_t*>(mine);
    if (ng && g0 + (ng - 1) * row + 4 <= ewords) {  // else a corrupt record: no AC entries
      for (uint32_t g = 0; g < ng; g += 4) {
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          v[k] = g + k < ng ? *reinterpret_cast<const uint4*>(ebuf + g0 + (g + k) * row)
                            : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t e[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
          for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t kk = 4 * (g + k) + q;
            if (kk >= ph && kk < ph + n && e[q] != 0xFFFFFFFFu) B[(e[q] >> 16) & 63] = (int16_t)(e[q] & 0xFFFFu);
          }
        }
      }
    }
    B[0] = r.dc;
  }
  __syncthreads();
  uint4* out = reinterpret_cast<uint4*>(coefs + b0 * 64);
  const uint64_t nq = (nblocks - b0 < 256 ? nblocks - b0 : 256) * 8;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t q = (uint32_t)(tid + 256 * i);  // 16-byte piece q & 7 of block q >> 3
    if (q < nq && !(recs[b0 + (q >> 3)].nw & kRecDense)) {
      const uint32_t* w = blk + (q >> 3) * kExpandStride + (q & 7) * 4;
      out[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

__global__ void k_entropy_serial(const uint32_t* __restrict__ words, const ImageDesc* __restrict__ images,
                                 const SegmentDesc* __restrict__ segs, uint32_t nsegs,
                                 const HuffTable* __restrict__ tables, int16_t* __restrict__ coefs,
                                 BlockRec* __restrict__ recs, uint32_t* __restrict__ status, uint32_t epoch) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsegs) return;
  ent_serial_segment(words, images, segs[s], tables, coefs, recs, status, epoch);
}

hipError_t launch_decode_begin(uint32_t* scratch, uint32_t nstatus, uint32_t nrec, hipStream_t s) {
  const uint32_t n = kScratchHead + nstatus + nrec;
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_decode_begin, dim3(blocks), dim3(256), 0, s, scratch, nstatus, nrec);
  return hipGetLastError();
}

hipError_t launch_clear(uint32_t* p, uint64_t nwords, hipStream_t s) {
  if (!nwords) return hipSuccess;
  uint64_t blocks = (nwords + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_clear_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, nwords);
  return hipGetLastError();
}

hipError_t launch_entropy(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                          const WorkItem* items, uint32_t item0, uint32_t item_end, const HuffTable* tables,
                          const uint16_t* luts, uint32_t* ebuf, BlockRec* recs, uint32_t* status, LookbackRec* lb,
                          unsigned long long* ticket, unsigned long long* done, uint64_t ticket_base, uint32_t epoch,
                          uint64_t wait_ticks, hipStream_t s) {
  if (item_end <= item0) return hipSuccess;
  hipLaunchKernelGGL(k_entropy, dim3(item_end - item0), dim3(kWG), 0, s, words, images, segs, items, item0, item_end,
                     tables, luts, ebuf, recs, status, lb, ticket, done, ticket_base, epoch, wait_ticks);
  return hipGetLastError();
}

hipError_t launch_expand(const BlockRec* recs, const uint32_t* ebuf, uint64_t ewords, int16_t* coefs, uint64_t nblocks,
                         hipStream_t s) {
  if (!nblocks) return hipSuccess;
  hipLaunchKernelGGL(k_expand, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, s, recs, ebuf, ewords, coefs,
                     nblocks);
  return hipGetLastError();
}

hipError_t launch_entropy_serial(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                                 uint32_t nsegs, const HuffTable* tables, int16_t* coefs, BlockRec* recs,
                                 uint32_t* status, uint32_t epoch, hipStream_t s) {
  if (!nsegs) return hipSuccess;
  hipLaunchKernelGGL(k_entropy_serial, dim3((nsegs + 63) / 64), dim3(64), 0, s, words, images, segs, nsegs, tables,
                     coefs, recs, status, epoch);
  return hipGetLastError();
}

}  // namespace cjd
