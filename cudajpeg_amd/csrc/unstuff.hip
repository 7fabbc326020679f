// Device-side unstuffing of entropy-coded scans (gfx950): the phases of
// unstuff_core.h, one 256-lane workgroup per chunk of 4096 stuffed bytes.
// Replaces the host loop of src/decodeScanGPU.cu:11-24 when the batch is
// staged with device unstuffing.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "unstuff_core.h"

namespace cjd {

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void __launch_bounds__(kUnstuffLanes) k_unstuff(const UnstuffChunk* __restrict__ chunks,
                                                           const uint8_t* __restrict__ raw,
                                                           const SegmentDesc* __restrict__ segs,
                                                           uint8_t* __restrict__ out) {
  __shared__ UnstuffShared sh;
  const int lane = threadIdx.x;
  const UnstuffChunk c = chunks[blockIdx.x];
  unstuff_load(sh, lane, c, raw);
  __syncthreads();
  if (lane == 0) unstuff_load_next(sh, c, raw);
  __syncthreads();
  unstuff_carry(sh, lane, c);
  __syncthreads();
  for (int d = 1; d < kUnstuffLanes; d <<= 1) {
    const uint8_t v = unstuff_fn_read(sh, lane, d);
    __syncthreads();
    sh.fn[lane] = v;
    __syncthreads();
  }
  const bool consumed = unstuff_carry_in(sh, lane, c);
  unstuff_count(sh, lane, c, consumed);
  __syncthreads();
  for (int d = 1; d < kUnstuffLanes; d <<= 1) {
    uint32_t brk, tail;
    unstuff_count_read(sh, lane, d, brk, tail);
    __syncthreads();
    sh.brk[lane] = brk;
    sh.tail[lane] = tail;
    __syncthreads();
  }
  if (unstuff_single_run(sh)) {
    // gathered in LDS, written as aligned words
    const uint64_t dst0 = segs[c.seg].word_base * 4u + c.dst;
    const uint32_t lead = (uint32_t)(dst0 & 15u);
    unstuff_gather(sh, lane, c, consumed, lead);
    __syncthreads();
    unstuff_flush(sh, lane, sh.tail[kUnstuffLanes - 1], lead, dst0, out);
  } else {
    unstuff_write(sh, lane, c, consumed, out, [&](uint32_t s) { return segs[s].word_base * 4u; });
  }
}

hipError_t launch_unstuff(const UnstuffChunk* chunks, uint32_t nchunks, const uint8_t* raw, const SegmentDesc* segs,
                          uint32_t* words, uint64_t nwords, hipStream_t s) {
  if (nwords) {
    uint64_t blocks = (nwords + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, words, nwords, 0xFFFFFFFFu);
  }
  if (nchunks)
    hipLaunchKernelGGL(k_unstuff, dim3(nchunks), dim3(kUnstuffLanes), 0, s, chunks, raw, segs,
                       reinterpret_cast<uint8_t*>(words));
  return hipGetLastError();
}

}  // namespace cjd
