// Fused pixel-stage kernel of the MI355X baseline-JPEG decoder (gfx950): the
// phases of pixel_core.h, one workgroup per run of MCUs.  Replaces iDCT_GPU and
// upsampleAndYUVtoRGB_kernel (src/pixelTransformGPU.cu:25-142, :475-511).
#include <hip/hip_runtime.h>

#include "launch.h"
#include "pixel_dev.h"

namespace cjd {

// Consecutive items per workgroup: 1 measured fastest (many small workgroups
// interleave best with an entropy launch running beside them).
constexpr int kPixItemsPerWG = 1;

// Table operands of the MFMA iDCT, computed once per image at staging: for
// each of the image's distinct quant tables t, K half h and part (hi, lo),
// lane tid = 64 n + l of the pixel kernel loads its 8 FP16 elements
// (mfma_b_lane for N tile n, lane l) from btab[image * kMfmaTabVecs +
// ((t * 2 + h) * 2 + part) * 256 + tid].
__global__ void __launch_bounds__(256) k_mfma_tables(const ImageDesc* __restrict__ images,
                                                     const uint16_t* __restrict__ qtabs, uint4* __restrict__ btab) {
  __shared__ uint16_t q[3][64];
  __shared__ float cosf[64];
  const int tid = threadIdx.x;
  const ImageDesc& im = images[blockIdx.x];
  for (int i = tid; i < im.ntab * 64; i += 256) q[i >> 6][i & 63] = qtabs[(im.qtab_base + im.tab_dq[i >> 6]) * 64 + (i & 63)];
  if (tid < 64) cosf[tid] = kCosF[tid >> 3][tid & 7];
  __syncthreads();
  uint4* dst = btab + (size_t)blockIdx.x * kMfmaTabVecs + tid;
  for (int t = 0; t < im.ntab; ++t)
    for (int h = 0; h < 2; ++h) {
      uint16_t hb[8], lb[8];
      mfma_b_lane(q[t], cosf, tid >> 6, h, tid & 63, hb, lb);
      uint4 vh, vl;
      __builtin_memcpy(&vh, hb, 16);
      __builtin_memcpy(&vl, lb, 16);
      dst[((t * 2 + h) * 2) * 256] = vh;
      dst[((t * 2 + h) * 2 + 1) * 256] = vl;
    }
}

// One workgroup decodes kPixItemsPerWG consecutive items.
template <int kMode>
__global__ void __launch_bounds__(kPixThreads) k_pixel(const ImageDesc* __restrict__ images,
                                                       const PixItem* __restrict__ items, uint32_t nitems,
                                                       const BlockRec* __restrict__ recs,
                                                       const uint32_t* __restrict__ ebuf, uint64_t ewords,
                                                       const int16_t* __restrict__ coefs,
                                                       const uint16_t* __restrict__ qtabs,
                                                       const uint4* __restrict__ btab,
                                                       uint8_t* __restrict__ out) {
  __shared__ PixSh<kMode> sh;
  const int tid = threadIdx.x;
  pix_stage_cos(sh, tid);
  MfmaB B;
  uint32_t b_image = ~0u;
  const PixArgs a{images, items, recs, ebuf, ewords, coefs, qtabs, btab, out};
  for (int k = 0; k < kPixItemsPerWG; ++k) {
    const uint32_t idx = blockIdx.x * kPixItemsPerWG + k;
    if (idx >= nitems) break;
    pixel_item<kMode>(sh, tid, idx, B, b_image, a);
  }
}

hipError_t launch_mfma_tables(const ImageDesc* images, uint32_t nimages, const uint16_t* qtabs, uint4* btab,
                              hipStream_t s) {
  if (!nimages) return hipSuccess;
  hipLaunchKernelGGL(k_mfma_tables, dim3(nimages), dim3(256), 0, s, images, qtabs, btab);
  return hipGetLastError();
}

hipError_t launch_pixel(int mode, const ImageDesc* images, const PixItem* items, uint32_t nitems,
                        const BlockRec* recs, const uint32_t* ebuf, uint64_t ewords, const int16_t* coefs,
                        const uint16_t* qtabs, const uint4* btab, uint8_t* out, hipStream_t s) {
  if (!nitems) return hipSuccess;
  const dim3 grid((nitems + kPixItemsPerWG - 1) / kPixItemsPerWG);
  switch (mode) {
    case kIdctInt:
      hipLaunchKernelGGL(k_pixel<kIdctInt>, grid, dim3(kPixThreads), 0, s, images, items, nitems, recs, ebuf, ewords,
                         coefs, qtabs, btab, out);
      break;
    case kIdctFloat:
      hipLaunchKernelGGL(k_pixel<kIdctFloat>, grid, dim3(kPixThreads), 0, s, images, items, nitems, recs, ebuf,
                         ewords, coefs, qtabs, btab, out);
      break;
    case kIdctMfma:
      hipLaunchKernelGGL(k_pixel<kIdctMfma>, grid, dim3(kPixThreads), 0, s, images, items, nitems, recs, ebuf, ewords,
                         coefs, qtabs, btab, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cjd
