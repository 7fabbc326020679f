// Entropy-decode kernels of the MI355X baseline-JPEG decoder (gfx950).
//
// k_entropy runs the phases of entropy_core.h for one work item (a run of
// tiles of one restart segment) per 256-lane workgroup, with a decoupled
// lookback across workgroups; k_entropy_serial is the one-lane-per-segment
// decoder used when the final decode disagrees with the chain.  Together they replace
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

__global__ void k_clear_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Fix-up rounds (all lanes): relink every tile whose entry changed, until no
// tile changes.
__device__ __forceinline__ void fix_rounds(EntShared& sh, int tid, const EntCtx& c) {
  for (;;) {
    uint32_t p = 0, s = 0;
    const bool need = ent_fix_need(sh, tid, c, p, s);
    if (tid == 0) sh.nchg = 0;
    __syncthreads();  // reads of the previous state done, counter reset
    if (need) {
      ent_fix_apply(sh, tid, c, p, s);
      atomicAdd(&sh.nchg, 1);
    }
    __syncthreads();
    const int n = sh.nchg;
    __syncthreads();  // every lane has read the counter before the next reset
    if (!n) break;
  }
}

// Inclusive prefix sum of the per-tile counts into tc (all lanes).
__device__ __forceinline__ void scan_counts(EntShared& sh, int tid, const EntCtx& c) {
  ent_scan_init(sh, tid, c);
  __syncthreads();
  for (int d = 1; d < c.J; d <<= 1) {
    int32_t v[4];
    ent_scan_read(sh, tid, d, v);
    __syncthreads();
    if (tid < c.J) ent_scan_write(sh, tid, v);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kWG) k_entropy(const uint32_t* __restrict__ words,
                                                 const ImageDesc* __restrict__ images,
                                                 const SegmentDesc* __restrict__ segs,
                                                 const WorkItem* __restrict__ items, uint32_t nitems,
                                                 const HuffTable* __restrict__ tables,
                                                 int16_t* __restrict__ coefs, uint32_t* __restrict__ status,
                                                 LookbackRec* __restrict__ lb, uint32_t* __restrict__ ticket) {
  __shared__ EntShared sh;
  const int tid = threadIdx.x;
  // Work items are taken in ticket order, so every predecessor an item waits
  // for in the lookback has already started.
  if (tid == 0) sh.item = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t item_id = sh.item;
  if (item_id >= nitems) return;
  EntCtx c;
  c.init(words, images, segs, items[item_id]);
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(tables + c.im->table_base);
    uint32_t* dst = reinterpret_cast<uint32_t*>(sh.tab);
    for (int i = tid; i < (int)(sizeof(HuffTable) * 4 / 4); i += kWG) dst[i] = src[i];
    if (tid < 64) sh.dz[tid] = kDeZigZag[tid];
  }
  __syncthreads();
  ent_phase1(sh, tid, c);
  __syncthreads();
  ent_phase2(sh, tid, c);
  __syncthreads();
  fix_rounds(sh, tid, c);
  scan_counts(sh, tid, c);
  LookbackRec* me = lb + item_id;
  if (tid == 0 && c.lead) {
    ent_aggregate(sh, c, me);
    st_status(&me->status, 1);
    // decoupled lookback: compose predecessors' aggregates until one has its
    // inclusive state
    LookAcc acc;
    acc.reset();
    uint32_t q_item = c.it.prev_item, p = 0, s = 0;
    int32_t cc[4] = {0, 0, 0, 0};
    for (;;) {
      const LookbackRec* r = lb + q_item;
      int st;
      while ((st = ld_status(&r->status)) == 0) __builtin_amdgcn_s_sleep(2);
      if (st == 2) {
        if (!acc.apply(*r, p, s, cc)) {
          // an assumption on the way does not hold: take the immediate
          // predecessor's inclusive state
          const LookbackRec* pr = lb + c.it.prev_item;
          while (ld_status(&pr->status) != 2) __builtin_amdgcn_s_sleep(2);
          acc.reset();
          acc.apply(*pr, p, s, cc);
        }
        break;
      }
      acc.absorb(*r);
      q_item = items[q_item].prev_item;
    }
    sh.entry_p = p;
    sh.entry_s = s;
    for (int q = 0; q < 4; ++q) sh.entry_c[q] = cc[q];
  }
  __syncthreads();
  if (ent_entry_differs(sh, c)) {  // uniform: the chain is redone from the true entry
    if (tid == 0) {
      sh.tp[0] = sh.entry_p;
      sh.ts[0] = sh.entry_s;
    }
    __syncthreads();
    fix_rounds(sh, tid, c);
    scan_counts(sh, tid, c);
  }
  ent_absolute(sh, tid, c);
  __syncthreads();
  if (tid == 0) {
    ent_inclusive(sh, c, me);
    st_status(&me->status, 2);
  }
  ent_phase5(sh, tid, c, coefs, status);
}

__global__ void k_entropy_serial(const uint32_t* __restrict__ words, const ImageDesc* __restrict__ images,
                                 const SegmentDesc* __restrict__ segs, uint32_t nsegs,
                                 const HuffTable* __restrict__ tables, int16_t* __restrict__ coefs,
                                 uint32_t* __restrict__ status) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsegs) return;
  ent_serial_segment(words, images, segs[s], tables, coefs, status);
}

hipError_t launch_clear(uint32_t* p, uint64_t nwords, hipStream_t s) {
  if (!nwords) return hipSuccess;
  uint64_t blocks = (nwords + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_clear_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, nwords);
  return hipGetLastError();
}

hipError_t launch_entropy(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                          const WorkItem* items, uint32_t nitems, const HuffTable* tables, int16_t* coefs,
                          uint32_t* status, LookbackRec* lb, uint32_t* ticket, hipStream_t s) {
  if (!nitems) return hipSuccess;
  hipLaunchKernelGGL(k_entropy, dim3(nitems), dim3(kWG), 0, s, words, images, segs, items, nitems, tables, coefs,
                     status, lb, ticket);
  return hipGetLastError();
}

hipError_t launch_entropy_serial(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                                 uint32_t nsegs, const HuffTable* tables, int16_t* coefs, uint32_t* status,
                                 hipStream_t s) {
  if (!nsegs) return hipSuccess;
  hipLaunchKernelGGL(k_entropy_serial, dim3((nsegs + 63) / 64), dim3(64), 0, s, words, images, segs, nsegs, tables,
                     coefs, status);
  return hipGetLastError();
}

}  // namespace cjd
