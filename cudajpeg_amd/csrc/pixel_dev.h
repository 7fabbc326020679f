// Device side of the fused pixel stage (gfx950): the MFMA iDCT of one item
// and one pixel work item's whole decode (coefficients from the block
// records, dequantisation, iDCT, planes, upsampling and colour) as a device
// function, called by k_pixel (pixel.hip) for each item it takes.
// Replaces iDCT_GPU and upsampleAndYUVtoRGB_kernel
// (src/pixelTransformGPU.cu:25-142, :475-511).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "launch.h"
#include "pixel_core.h"

namespace cjd {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// B operands of one wave: [table slot][K half], hi and lo parts.
struct MfmaB {
  half8_t hi[2][2], lo[2][2];
};

// The 8 products of a tile pair whose blocks all use table slot TL: per K
// half, hi and lo parts, the two tiles interleaved.
template <int TL>
__device__ __forceinline__ void mfma_one_table(const MfmaB& B, const half8_t (&a)[2][2], float4_t (&acc)[2]) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(B.hi[TL][h], a[jj][h], acc[jj], 0, 0, 0);
      acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(B.lo[TL][h], a[jj][h], acc[jj], 0, 0, 0);
    }
}

// Batched-form iDCT of one item on the matrix cores (see the MFMA section of
// pixel_core.h).  The contraction runs as C[pixel][block] = M_t^T x coef^T:
// wave w computes N tile w (pixels 16w .. 16w+15) of every 16-block tile, so
// its matrix operands depend only on the image's quant tables and are built
// in registers (loaded from the image's k_mfma_tables output) and kept while
// consecutive items belong to the same image (`b_image`, ~0 = none); lane l
// receives 4 consecutive pixels (one 4-byte
// sample word) of block l % 16.  The item's blocks sit in LDS in table order,
// and tiles go two at a time: a pair whose blocks all use one table runs that
// table's 8 products only; a pair across a table boundary runs every table of
// the pass with the blocks of the other tables masked to zero.  All LDS reads
// of a pair are issued together from clamped, always-valid rows.
__device__ __forceinline__ void pix_idct_mfma(PixSharedMfma& sh, int tid, const PixCtx& c, MfmaB& B,
                                              uint32_t& b_image, const uint4* __restrict__ btab) {
  const int n = tid >> 6, l = tid & 63;
  const int ntiles = (c.nblk + 15) >> 4;
  const int last = c.nblk - 1;
  const half8_t zero = {0, 0, 0, 0, 0, 0, 0, 0};
  const int kofs = 8 * (l >> 4);
  const int yo = 2 * n + (l >> 5), xo = 4 * ((l >> 4) & 1);  // the lane's pixel row and first column
  for (int t0 = 0; t0 < c.ntab; t0 += 2) {
    const int nt = c.ntab - t0 < 2 ? c.ntab - t0 : 2;
    if (t0 > 0 || b_image != c.it.image) {
      // constant loop bounds keep the operand arrays in registers
      const uint4* src = btab + (size_t)c.it.image * kMfmaTabVecs + tid;
      for (int tl = 0; tl < 2; ++tl)
        for (int h = 0; h < 2; ++h) {
          if (tl >= nt) continue;
          const uint4 vh = src[(((t0 + tl) * 2 + h) * 2) * 256], vl = src[(((t0 + tl) * 2 + h) * 2 + 1) * 256];
          __builtin_memcpy(&B.hi[tl][h], &vh, 16);
          __builtin_memcpy(&B.lo[tl][h], &vl, 16);
        }
    }
    for (int tile = 0; tile < ntiles; tile += 2) {
      half8_t a[2][2];
      int tix[2];
      uint32_t info[2];
      for (int jj = 0; jj < 2; ++jj) {
        const int pj = (tile + jj) * 16 + (l & 15);
        const int r = pj <= last ? pj : last;
        __builtin_memcpy(&a[jj][0], &sh.coef[r][kofs], 16);
        __builtin_memcpy(&a[jj][1], &sh.coef[r][32 + kofs], 16);
        info[jj] = sh.binfo[r];
        tix[jj] = pj <= last ? binfo_tix(info[jj]) : -1;
      }
      const int f = tile * 16, e = f + 31 <= last ? f + 31 : last;
      const int tlo = __builtin_amdgcn_readfirstlane(binfo_tix(sh.binfo[f])) - t0;
      const int thi = __builtin_amdgcn_readfirstlane(binfo_tix(sh.binfo[e])) - t0;
      float4_t acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
      if (tlo == thi) {
        if (tlo == 0)
          mfma_one_table<0>(B, a, acc);
        else if (tlo == 1 && nt > 1)
          mfma_one_table<1>(B, a, acc);
      } else {
        for (int tl = 0; tl < 2; ++tl)
          for (int h = 0; h < 2; ++h)
            for (int jj = 0; jj < 2; ++jj) {
              if (tl >= nt) continue;
              const half8_t bm = tix[jj] == t0 + tl ? a[jj][h] : zero;
              acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(B.hi[tl][h], bm, acc[jj], 0, 0, 0);
              acc[jj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(B.lo[tl][h], bm, acc[jj], 0, 0, 0);
            }
      }
      for (int jj = 0; jj < 2; ++jj)
        if (tix[jj] >= t0 && tix[jj] < t0 + nt) {
          uint32_t word = 0;
          for (int i = 0; i < 4; ++i) word |= (uint32_t)clip8((int)floorf(acc[jj][i] + 128.5f)) << (8 * i);
          *reinterpret_cast<uint32_t*>(&sh.samp[binfo_base(info[jj]) + yo * binfo_tw(info[jj]) + xo]) = word;
        }
    }
  }
  b_image = c.ntab <= 2 ? c.it.image : ~0u;
}


template <int kMode>
using PixSh = typename std::conditional<kMode == kIdctMfma, PixSharedMfma, PixShared>::type;

// Pixel work item `idx` by the calling workgroup (all lanes).  Barriers:
// previous item done -> zeroed blocks -> block fill -> iDCT (coefficients,
// descriptors, quant tables) -> colour (samples).  The caller has staged the
// cosine table (pix_stage_cos) once; B and b_image carry the MFMA table
// operands from item to item of the same image.
template <int kMode>
__device__ inline void pixel_item(PixSh<kMode>& sh, int tid, uint32_t idx, MfmaB& B, uint32_t& b_image,
                                  const PixArgs& a) {
  PixCtx c;
  c.init(a.images, a.items[idx]);
  __syncthreads();  // the previous item is done with coef, samp, q and binfo
  pix_stage_zero(sh, tid, c, a.qtabs);
  __syncthreads();
  pix_stage_blocks(sh, tid, c, a.recs, a.ebuf, a.ewords, a.coefs);
  __syncthreads();
  if constexpr (kMode == kIdctMfma) {
    pix_idct_mfma(sh, tid, c, B, b_image, a.btab);
  } else {
    pix_rows<kMode>(sh, tid, c);
    __syncthreads();
    pix_cols<kMode>(sh, tid, c);
  }
  __syncthreads();
  pix_planes(sh, tid, c, a.out);
  pix_colour(sh, tid, c, a.out);
}

}  // namespace cjd
