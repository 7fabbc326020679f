// Fused pixel-stage kernel of the MI355X baseline-JPEG decoder (gfx950): the
// phases of pixel_core.h, one workgroup per run of MCUs.  Replaces iDCT_GPU and
// upsampleAndYUVtoRGB_kernel (src/pixelTransformGPU.cu:25-142, :475-511).
#include <hip/hip_runtime.h>

#include "launch.h"
#include "pixel_core.h"

namespace cjd {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Batched-form iDCT on the matrix cores: wave w takes 16-block tile w of the
// workgroup and runs one pass per distinct quant table (see pixel_core.h).
__device__ __forceinline__ void pix_idct_mfma(PixShared& sh, int tid, const PixCtx& c) {
  const int wave = tid >> 6, l = tid & 63;
  const int ntiles = (c.nblk + 15) >> 4;
  if (wave >= ntiles) return;
  const int b0 = wave * 16;
  const int brow = b0 + (l & 15);              // block of this lane's A row
  const int bcol0 = b0 + 4 * (l >> 4);         // first block of this lane's C rows
  int tab_row = -1;
  if (brow < c.nblk) tab_row = c.im->comp_dq[c.im->slot_comp[brow % c.nslots]];
  int tab_out[4];
  for (int i = 0; i < 4; ++i) {
    const int b = bcol0 + i;
    tab_out[i] = b < c.nblk ? c.im->comp_dq[c.im->slot_comp[b % c.nslots]] : -1;
  }
  // A operands (coefficients, exact in FP16), loaded once
  half8_t a[2];
  for (int kk = 0; kk < 2; ++kk) {
    int16_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (brow < c.nblk) __builtin_memcpy(v, &sh.coef[brow][kk * 32 + 8 * (l >> 4)], 16);
    for (int i = 0; i < 8; ++i) a[kk][i] = (_Float16)v[i];
  }
  // one pass per distinct quant table among the components
  for (int comp = 0; comp < c.ncomp; ++comp) {
    const int t = c.im->comp_dq[comp];
    bool seen = false;
    for (int e = 0; e < comp; ++e) seen |= c.im->comp_dq[e] == t;
    if (seen) continue;
    float hi[2][4][8], lo[2][4][8];
    mfma_basis(sh.q[comp], l, hi, lo);
    half8_t am[2];
    for (int kk = 0; kk < 2; ++kk)
      for (int i = 0; i < 8; ++i) am[kk][i] = tab_row == t ? a[kk][i] : (_Float16)0.0f;
    for (int n = 0; n < 4; ++n) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
      for (int kk = 0; kk < 2; ++kk) {
        half8_t bh, bl;
        for (int i = 0; i < 8; ++i) {
          bh[i] = (_Float16)hi[kk][n][i];
          bl[i] = (_Float16)lo[kk][n][i];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(am[kk], bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(am[kk], bl, acc, 0, 0, 0);
      }
      const int p = n * 16 + (l & 15);
      for (int i = 0; i < 4; ++i)
        if (tab_out[i] == t) mfma_store(sh, c, bcol0 + i, p, acc[i]);
    }
  }
}

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
  if (kMode == kIdctMfma) {
    pix_idct_mfma(sh, tid, c);
  } else {
    pix_rows<kMode>(sh, tid, c);
    __syncthreads();
    pix_cols<kMode>(sh, tid, c);
  }
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
    case kIdctMfma:
      hipLaunchKernelGGL(k_pixel<kIdctMfma>, dim3(nitems), dim3(kPixThreads), 0, s, images, items, coefs, qtabs,
                         out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace cjd
