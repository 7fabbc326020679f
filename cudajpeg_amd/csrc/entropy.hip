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

#include "entropy_dev.h"
#include "launch.h"

namespace cjd {

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
  if (tid == 0) sh.item = item0 + (uint32_t)(atomicAdd(ticket, 1ull) - ticket_base);
  __syncthreads();
  const uint32_t item_id = sh.item;
  if (item_id < item0 || item_id >= item_end) return;
  const EntArgs a{words, images, segs, items, tables, luts, ebuf, recs, status, lb, done, epoch, wait_ticks};
  entropy_item(sh, tid, item_id, a);
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
