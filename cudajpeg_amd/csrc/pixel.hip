// Fused pixel-stage kernel of the MI355X baseline-JPEG decoder (gfx950): the
// phases of pixel_core.h, one workgroup per run of MCUs.  Replaces iDCT_GPU and
// upsampleAndYUVtoRGB_kernel (src/pixelTransformGPU.cu:25-142, :475-511).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "launch.h"
#include "pixel_core.h"

namespace cjd {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Consecutive items per workgroup: 1 measured fastest (many small workgroups
// interleave best with an entropy launch running beside them).
constexpr int kPixItemsPerWG = 1;

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

// One workgroup decodes kPixItemsPerWG consecutive items.  Barriers: zeroed
// blocks -> block fill; staging -> iDCT (coefficients, descriptors, quant
// tables); iDCT -> colour (samples); colour -> next item's staging.
template <int kMode>
__global__ void __launch_bounds__(kPixThreads) k_pixel(const ImageDesc* __restrict__ images,
                                                       const PixItem* __restrict__ items, uint32_t nitems,
                                                       const BlockRec* __restrict__ recs,
                                                       const uint32_t* __restrict__ ebuf, uint64_t ewords,
                                                       const int16_t* __restrict__ coefs,
                                                       const uint16_t* __restrict__ qtabs,
                                                       const uint4* __restrict__ btab,
                                                       uint8_t* __restrict__ out) {
  using SH = typename std::conditional<kMode == kIdctMfma, PixSharedMfma, PixShared>::type;
  __shared__ SH sh;
  const int tid = threadIdx.x;
  pix_stage_cos(sh, tid);
  MfmaB B;
  uint32_t b_image = ~0u;
  for (int k = 0; k < kPixItemsPerWG; ++k) {
    const uint32_t idx = blockIdx.x * kPixItemsPerWG + k;
    if (idx >= nitems) break;
    PixCtx c;
    c.init(images, items[idx]);
    __syncthreads();  // the previous item is done with coef, samp, q and binfo
    pix_stage_zero(sh, tid, c, qtabs);
    __syncthreads();
    pix_stage_blocks(sh, tid, c, recs, ebuf, ewords, coefs);
    __syncthreads();
    if constexpr (kMode == kIdctMfma) {
      pix_idct_mfma(sh, tid, c, B, b_image, btab);
    } else {
      pix_rows<kMode>(sh, tid, c);
      __syncthreads();
      pix_cols<kMode>(sh, tid, c);
    }
    __syncthreads();
    pix_planes(sh, tid, c, out);
    pix_colour(sh, tid, c, out);
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
