// Progressive-scan kernels (CJD_EXT_PROGRESSIVE); the decode itself is in
// prog_core.h.  Beyond the reference, which rejects SOF2
// (src/format.cu:198-200).
#include <hip/hip_runtime.h>

#include "launch.h"
#include "prog_core.h"

namespace cjd {

// One thread per block of the progressive images: coefficients zeroed, a
// dense record.
__global__ void __launch_bounds__(256) k_prog_prepare(const ProgImage* __restrict__ pimgs, uint32_t n, int64_t total,
                                                      int16_t* __restrict__ coefs, BlockRec* __restrict__ recs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) prog_prepare_block(pimgs, n, i, coefs, recs);
}

// One wave64 per task of a level (prog_core.h: the lanes decode in lockstep,
// lane z owning the coefficient at zigzag position z).
__global__ void __launch_bounds__(64) k_prog(const ProgTask* __restrict__ tasks, uint32_t first,
                                             const uint32_t* __restrict__ words, const ImageDesc* __restrict__ images,
                                             const SegmentDesc* __restrict__ segs,
                                             const HuffTable* __restrict__ tables, int16_t* __restrict__ coefs,
                                             uint32_t* __restrict__ status, uint32_t epoch) {
  __shared__ ProgShared sh;
  const ProgTask t = tasks[first + blockIdx.x];
  const SegmentDesc sg = segs[t.seg];
  prog_task(sh, words, images, sg, t, tables + (size_t)sg.tset * 4, coefs, status, epoch);
}

hipError_t launch_prog(const ProgImage* pimgs, uint32_t npimgs, int64_t prog_blocks, const ProgTask* tasks,
                       const uint32_t* plevel, uint32_t nlevels, const uint32_t* words, const ImageDesc* images,
                       const SegmentDesc* segs, const HuffTable* tables, int16_t* coefs, BlockRec* recs,
                       uint32_t* status, uint32_t epoch, hipStream_t s) {
  if (prog_blocks <= 0) return hipSuccess;
  const uint64_t g = ((uint64_t)prog_blocks + 255) / 256;
  if (g > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_prog_prepare, dim3((uint32_t)g), dim3(256), 0, s, pimgs, npimgs, prog_blocks, coefs, recs);
  hipError_t e = hipGetLastError();
  for (uint32_t l = 0; l < nlevels && e == hipSuccess; ++l) {
    const uint32_t count = plevel[l + 1] - plevel[l];
    if (!count) continue;
    hipLaunchKernelGGL(k_prog, dim3(count), dim3(64), 0, s, tasks, plevel[l], words, images, segs, tables, coefs,
                       status, epoch);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace cjd
