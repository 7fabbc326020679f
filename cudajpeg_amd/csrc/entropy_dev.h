// Device side of the tiled entropy decoder (gfx950): the wave64 primitives,
// the workgroup orchestration of the phases of entropy_core.h, and one work
// item's whole decode as a device function, called by k_entropy (one item
// per workgroup, entropy.hip).  Replaces the
// speculative kernels of src/entropyRLEdecodeGPU.cu:7-238 and the host
// serial decode of src/decodeScanGPU.cu:435-451.
#pragma once
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

// The decode of work item `item_id` by the calling workgroup (all lanes):
// lookup tables into LDS, speculative pass, links, fix-up rounds, counts and
// their scan, the decoupled lookback, and the block records.  Every
// predecessor the lookback waits for must already be running (items are
// taken in ticket order).  Ends with a barrier-free return on every path;
// the caller separates items with barriers.
__device__ inline void entropy_item(EntShared& sh, int tid, uint32_t item_id, const EntArgs& a) {
  if (tid == 0) {
    sh.lovf = 0;
    sh.look_ok = 1;
  }
  EntCtx c;
  c.init(a.words, a.images, a.segs, a.tables, a.ebuf, a.items[item_id]);
  {
    const uint4* src = reinterpret_cast<const uint4*>(a.luts + (size_t)c.sg.tset * kLutEntries);
    uint4* dst = reinterpret_cast<uint4*>(sh.lut);
    for (int i = tid; i < kLutEntries / 8; i += kWG) dst[i] = src[i];
    if (tid < 64) sh.dz[tid] = kDeZigZag[tid];
    if (tid < 4) canon_fill(sh.canon[tid], c.tabs[tid]);
    if (tid < c.nh) {
      sh.tsel[2 * tid] = (uint16_t)(seg_ac(c.sg, *c.im, tid) * kL1Size);
      sh.tsel[2 * tid + 1] = (uint16_t)(seg_dc(c.sg, *c.im, tid) * kL1Size);
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
  LookbackRec* me = a.lb + item_id;
  LookbackRec mine;
  if (tid == 0 && c.lead) {
    ent_aggregate(sh, c, mine);
    st_record_words(me, mine, kRecAgg0, kRecAgg1);
    st_status(&me->status, lookback_word(a.epoch, kLookAgg));
    // decoupled lookback: compose predecessors' aggregates until one has its
    // inclusive state.  A record of an older launch does not match this
    // epoch; every wait ends, after wait_ticks of the real-time counter at
    // most, and then the item gives up.
    auto wait = [&](uint32_t q, bool incl) -> int {
      const int32_t* w = &a.lb[q].status;
      uint64_t t0 = 0;
      for (bool first = true;; first = false) {
        const int st = ld_status(w);
        if (((uint32_t)st >> 2) == a.epoch && (!incl || (st & 3) >= kLookIncl)) return st;
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (first) t0 = now;
        else if (now - t0 >= a.wait_ticks) return -1;
        __builtin_amdgcn_s_sleep(2);
      }
    };
    uint32_t p = 0, s = 0;
    int32_t cc[4] = {0, 0, 0, 0};
    auto load = [&](uint32_t q) { return ld_record(a.lb + q); };
    if (lookback_walk(a.items, c.sg.first_item, item_id, c.it.prev_item, wait, load, p, s, cc)) {
      sh.entry_p = p;
      sh.entry_s = s;
      for (int q = 0; q < 4; ++q) sh.entry_c[q] = cc[q];
    } else {
      // successors give up at this record; the host decodes the image again
      sh.look_ok = 0;
      status_or(a.status + c.im->error_index, kStatusUnresolved, a.epoch);
      st_status(&me->status, lookback_word(a.epoch, kLookFail));
    }
  }
  __syncthreads();
  if (!sh.look_ok) {
    if (tid == 0) atomicAdd(a.done, 1ull);
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
    st_status(&me->status, lookback_word(a.epoch, kLookIncl));
    atomicAdd(a.done, 1ull);  // read by the host after the launch
  }
  ent_records(sh, tid, c, a.recs, a.status, a.epoch);
}

}  // namespace cjd
