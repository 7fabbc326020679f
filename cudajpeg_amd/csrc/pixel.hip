// Fused pixel-stage kernel of the MI355X baseline-JPEG decoder (gfx950): the
// phases of pixel_core.h, one workgroup per run of MCUs.  Replaces iDCT_GPU and
// upsampleAndYUVtoRGB_kernel (src/pixelTransformGPU.cu:25-142, :475-511).
#include <hip/hip_runtime.h>

#include "launch.h"
#include "pixel_core.h"

namespace cjd {

template <int kMode>
__global__ void __launch_bounds__(kPixThreads) k_pixel(const ImageDesc* __restrict__ images,
                                                       const PixItem* __restrict__ items,
                                                       const int16_t* __restrict__ coefs,
                                                       const uint16_t* __restrict__ qtabs,
                                                       uint8_t* __restrict__ out) {
  __shared__ PixShared sh;
  const int tid = threadIdx.x;
  PixCtx c;
  c.init(images, items[blockIdx.x]);
  pix_stage(sh, tid, c, coefs, qtabs);
  __syncthreads();
  pix_rows<kMode>(sh, tid, c);
  __syncthreads();
  pix_cols<kMode>(sh, tid, c);
  __syncthreads();
  pix_planes(sh, tid, c, out);
  pix_colour(sh, tid, c, out);
}

hipError_t launch_pixel(int mode, const ImageDesc* images, const PixItem* items, uint32_t nitems,
                        const int16_t* coefs, const uint16_t* qtabs, uint8_t* out, hipStream_t s) {
  if (!nitems) return hipSuccess;
  switch (mode) {
    case kIdctInt:
      hipLaunchKernelGGL(k_pixel<kIdctInt>, dim3(nitems), dim3(kPixThreads), 0, s, images, items, coefs, qtabs, out);
      break;
    case kIdctFloat:
      hipLaunchKernelGGL(k_pixel<kIdctFloat>, dim3(nitems), dim3(kPixThreads), 0, s, images, items, coefs, qtabs,
                         out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cjd
