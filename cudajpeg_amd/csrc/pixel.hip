// Fused pixel stage: dequant + 8x8 iDCT + nearest chroma upsample + YCbCr->RGB
// (or gray copy), one write per output pixel.
//
// Replaces src/pixelTransformGPU.cu:25-142 (iDCT_GPU), :475-511
// (upsampleAndYUVtoRGB_kernel) and :621-677 (iDCT_resample_colourTransform),
// with the edge semantics of CPUonly/src/pixelTransformCPU.c:110-172.
//
// One workgroup takes a run of consecutive MCUs of one MCU row.  Their
// quantised coefficients (contiguous in scan order) are staged in LDS, each
// block is inverse-transformed by 8 lanes (rows, then columns) into per-
// component sample tiles in LDS, and the tile is colour-converted straight to
// the interleaved output.  iDCT variants:
//   kIdctInt   : the reference's integer Chen-Wang iDCT, bit-exact
//                (CPUonly/src/pixelTransformCPU.c:14-107)
//   kIdctFloat : separable FP32 iDCT of the dequantised coefficients (T.81 A.3.3),
//                rounded to nearest
//   kIdctMfma  : 64x64 contraction on the matrix cores (FP16 hi/lo basis with the
//                quant table folded in, FP32 accumulate)
#include <hip/hip_runtime.h>

#include "cjd_internal.h"

namespace cjd {

// CPUonly/include/pixelTransformCPU.h:5-10
#define W1 2841
#define W2 2676
#define W3 2408
#define W5 1609
#define W6 1108
#define W7 565

__device__ __forceinline__ uint8_t clip8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

// Row pass, CPUonly/src/pixelTransformCPU.c:14-57 without the all-zero
// shortcut (which yields identical values).
__device__ __forceinline__ void idct_row_int(const int* D, int* O) {
  int x0, x1, x2, x3, x4, x5, x6, x7, x8;
  x1 = D[4] << 11; x2 = D[6]; x3 = D[2]; x4 = D[1]; x5 = D[7]; x6 = D[5]; x7 = D[3];
  x0 = (D[0] << 11) + 128;
  x8 = W7 * (x4 + x5);
  x4 = x8 + (W1 - W7) * x4;
  x5 = x8 - (W1 + W7) * x5;
  x8 = W3 * (x6 + x7);
  x6 = x8 - (W3 - W5) * x6;
  x7 = x8 - (W3 + W5) * x7;
  x8 = x0 + x1; x0 -= x1;
  x1 = W6 * (x3 + x2);
  x2 = x1 - (W2 + W6) * x2;
  x3 = x1 + (W2 - W6) * x3;
  x1 = x4 + x6; x4 -= x6; x6 = x5 + x7; x5 -= x7;
  x7 = x8 + x3; x8 -= x3; x3 = x0 + x2; x0 -= x2;
  x2 = (181 * (x4 + x5) + 128) >> 8;
  x4 = (181 * (x4 - x5) + 128) >> 8;
  O[0] = (x7 + x1) >> 8; O[1] = (x3 + x2) >> 8; O[2] = (x0 + x4) >> 8; O[3] = (x8 + x6) >> 8;
  O[4] = (x8 - x6) >> 8; O[5] = (x0 - x4) >> 8; O[6] = (x3 - x2) >> 8; O[7] = (x7 - x1) >> 8;
}

// Column pass, CPUonly/src/pixelTransformCPU.c:60-107.
__device__ __forceinline__ void idct_col_int(const int* D, uint8_t* out) {
  int x0, x1, x2, x3, x4, x5, x6, x7, x8;
  x1 = D[4] << 8; x2 = D[6]; x3 = D[2]; x4 = D[1]; x5 = D[7]; x6 = D[5]; x7 = D[3];
  x0 = (D[0] << 8) + 8192;
  x8 = W7 * (x4 + x5) + 4;
  x4 = (x8 + (W1 - W7) * x4) >> 3;
  x5 = (x8 - (W1 + W7) * x5) >> 3;
  x8 = W3 * (x6 + x7) + 4;
  x6 = (x8 - (W3 - W5) * x6) >> 3;
  x7 = (x8 - (W3 + W5) * x7) >> 3;
  x8 = x0 + x1; x0 -= x1;
  x1 = W6 * (x3 + x2) + 4;
  x2 = (x1 - (W2 + W6) * x2) >> 3;
  x3 = (x1 + (W2 - W6) * x3) >> 3;
  x1 = x4 + x6; x4 -= x6; x6 = x5 + x7; x5 -= x7;
  x7 = x8 + x3; x8 -= x3; x3 = x0 + x2; x0 -= x2;
  x2 = (181 * (x4 + x5) + 128) >> 8;
  x4 = (181 * (x4 - x5) + 128) >> 8;
  out[0] = clip8(((x7 + x1) >> 14) + 128);
  out[1] = clip8(((x3 + x2) >> 14) + 128);
  out[2] = clip8(((x0 + x4) >> 14) + 128);
  out[3] = clip8(((x8 + x6) >> 14) + 128);
  out[4] = clip8(((x8 - x6) >> 14) + 128);
  out[5] = clip8(((x0 - x4) >> 14) + 128);
  out[6] = clip8(((x3 - x2) >> 14) + 128);
  out[7] = clip8(((x7 - x1) >> 14) + 128);
}

// cos basis for the FP32 path: kCosF[u][x] = C(u)/2 cos((2x+1) u pi / 16)
__constant__ float kCosF[8][8] = {
    {0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f},
    {0.49039264f, 0.41573481f, 0.27778512f, 0.09754516f, -0.09754516f, -0.27778512f, -0.41573481f, -0.49039264f},
    {0.46193977f, 0.19134172f, -0.19134172f, -0.46193977f, -0.46193977f, -0.19134172f, 0.19134172f, 0.46193977f},
    {0.41573481f, -0.09754516f, -0.49039264f, -0.27778512f, 0.27778512f, 0.49039264f, 0.09754516f, -0.41573481f},
    {0.35355339f, -0.35355339f, -0.35355339f, 0.35355339f, 0.35355339f, -0.35355339f, -0.35355339f, 0.35355339f},
    {0.27778512f, -0.49039264f, 0.09754516f, 0.41573481f, -0.41573481f, -0.09754516f, 0.49039264f, -0.27778512f},
    {0.19134172f, -0.46193977f, 0.46193977f, -0.19134172f, -0.19134172f, 0.46193977f, -0.46193977f, 0.19134172f},
    {0.09754516f, -0.27778512f, 0.41573481f, -0.49039264f, 0.49039264f, -0.41573481f, 0.27778512f, -0.09754516f}};

struct PixShared {
  int16_t coef[kPixBlocks][64];           // quantised coefficients, natural order
  union {
    int32_t ti[kPixBlocks][64];           // row-pass output (integer path)
    float tf[kPixBlocks][64];             // row-pass output (float path)
  } t;
  uint8_t samp[kPixBlocks * 64];          // component sample tiles
  uint16_t q[3][64];                      // natural-order quant table per component
};

template <int kMode>
__global__ void __launch_bounds__(256) k_pixel(const ImageDesc* __restrict__ images,
                                               const PixItem* __restrict__ items,
                                               const int16_t* __restrict__ coefs,
                                               const uint16_t* __restrict__ qtabs,
                                               uint8_t* __restrict__ out) {
  __shared__ PixShared sh;
  const PixItem it = items[blockIdx.x];
  const ImageDesc& im = images[it.image];
  const int tid = threadIdx.x;
  const int nslots = im.nslots;
  const int nblk = (int)it.count * nslots;
  const int ncomp = im.ncomp;

  // stage coefficients (contiguous: MCUs in scan order)
  {
    const uint4* src = reinterpret_cast<const uint4*>(coefs + (im.coef_base + (int64_t)it.mcu_first * nslots) * 64);
    uint4* dst = reinterpret_cast<uint4*>(&sh.coef[0][0]);
    for (int i = tid; i < nblk * 8; i += 256) dst[i] = src[i];
    for (int i = tid; i < ncomp * 64; i += 256) {
      const int c = i >> 6;
      sh.q[c][i & 63] = qtabs[(im.qtab_base + im.comp_dq[c]) * 64 + (i & 63)];
    }
  }
  __syncthreads();

  // component tile geometry: component c's tile is (count*sx*8) wide, sy*8 tall
  int toff[3], tw[3];
  {
    int off = 0;
    for (int c = 0; c < ncomp; ++c) {
      toff[c] = off;
      tw[c] = (int)it.count * im.comp_sx[c] * 8;
      off += tw[c] * im.comp_sy[c] * 8;
    }
  }

  // row pass: lane -> (block, row)
  for (int task = tid; task < nblk * 8; task += 256) {
    const int b = task >> 3, r = task & 7;
    const int slot = b % nslots;
    const int c = im.slot_comp[slot];
    const int16_t* cb = &sh.coef[b][r * 8];
    const uint16_t* qq = &sh.q[c][r * 8];
    if (kMode == kIdctInt) {
      int D[8], O[8];
      for (int k = 0; k < 8; ++k) D[k] = (int)cb[k] * (int)qq[k];
      idct_row_int(D, O);
      for (int k = 0; k < 8; ++k) sh.t.ti[b][r * 8 + k] = O[k];
    } else {
      float D[8];
      for (int k = 0; k < 8; ++k) D[k] = (float)((int)cb[k] * (int)qq[k]);
      for (int x = 0; x < 8; ++x) {
        float s = 0.f;
        for (int u = 0; u < 8; ++u) s = fmaf(kCosF[u][x], D[u], s);
        sh.t.tf[b][r * 8 + x] = s;
      }
    }
  }
  __syncthreads();

  // column pass: lane -> (block, column); samples into the component tile
  for (int task = tid; task < nblk * 8; task += 256) {
    const int b = task >> 3, x = task & 7;
    const int m = b / nslots, slot = b - m * nslots;
    const int c = im.slot_comp[slot];
    const int k = slot - im.comp_first_slot[c];
    const int sx = k % im.comp_sx[c], sy = k / im.comp_sx[c];
    uint8_t* dst = &sh.samp[toff[c] + (sy * 8) * tw[c] + (m * im.comp_sx[c] + sx) * 8 + x];
    if (kMode == kIdctInt) {
      int D[8];
      uint8_t o[8];
      for (int v = 0; v < 8; ++v) D[v] = sh.t.ti[b][v * 8 + x];
      idct_col_int(D, o);
      for (int y = 0; y < 8; ++y) dst[y * tw[c]] = o[y];
    } else {
      float D[8];
      for (int v = 0; v < 8; ++v) D[v] = sh.t.tf[b][v * 8 + x];
      for (int y = 0; y < 8; ++y) {
        float s = 128.f;
        for (int v = 0; v < 8; ++v) s = fmaf(kCosF[v][y], D[v], s);
        dst[y * tw[c]] = clip8((int)floorf(s + 0.5f));
      }
    }
  }
  __syncthreads();

  // optional planar outputs (reference plane layout, src/format.cu:295-303)
  const int mcu_y = (int)(it.mcu_first / im.mcus_x), mcu_x0 = (int)(it.mcu_first % im.mcus_x);
  for (int c = 0; c < ncomp; ++c) {
    if (im.plane_offset[c] == ~0ull) continue;
    uint8_t* pl = out + im.plane_offset[c];
    const int rows = im.comp_sy[c] * 8, w = tw[c];
    const int px0 = mcu_x0 * im.comp_sx[c] * 8, py0 = mcu_y * rows;
    for (int i = tid; i < rows * w; i += 256) {
      const int yy = i / w, xx = i - yy * w;
      pl[(size_t)(py0 + yy) * im.comp_stride[c] + px0 + xx] = sh.samp[toff[c] + yy * w + xx];
    }
  }

  // colour conversion / gray copy, one byte per lane per step
  // (CPUonly/src/pixelTransformCPU.c:130-172)
  const int x0 = mcu_x0 * im.mcu_w, y0 = mcu_y * im.mcu_h;
  int x1 = x0 + (int)it.count * im.mcu_w;
  if (x1 > im.width) x1 = im.width;
  int y1 = y0 + im.mcu_h;
  if (y1 > im.height) y1 = im.height;
  const int rw = x1 - x0, rows = y1 - y0;
  uint8_t* o = out + im.out_offset;
  if (ncomp == 1) {
    for (int i = tid; i < rw * rows; i += 256) {
      const int yy = i / rw, xx = i - yy * rw;
      o[(size_t)(y0 + yy) * im.width + x0 + xx] = sh.samp[yy * tw[0] + xx];
    }
  } else {
    const int xs0 = im.comp_xshift[0], ys0 = im.comp_yshift[0];
    const int xs1 = im.comp_xshift[1], ys1 = im.comp_yshift[1];
    const int xs2 = im.comp_xshift[2], ys2 = im.comp_yshift[2];
    const int nbytes = rw * 3;
    for (int i = tid; i < nbytes * rows; i += 256) {
      const int yy = i / nbytes, bx = i - yy * nbytes;
      const int xx = bx / 3, ch = bx - xx * 3;
      const int Y = (int)sh.samp[toff[0] + (yy >> ys0) * tw[0] + (xx >> xs0)] << 8;
      const int cb = (int)sh.samp[toff[1] + (yy >> ys1) * tw[1] + (xx >> xs1)] - 128;
      const int cr = (int)sh.samp[toff[2] + (yy >> ys2) * tw[2] + (xx >> xs2)] - 128;
      int v;
      if (ch == 0) v = (Y + 359 * cr + 128) >> 8;
      else if (ch == 1) v = (Y - 88 * cb - 183 * cr + 128) >> 8;
      else v = (Y + 454 * cb + 128) >> 8;
      o[((size_t)(y0 + yy) * im.width + x0) * 3 + bx] = clip8(v);
    }
  }
}

template __global__ void k_pixel<kIdctInt>(const ImageDesc*, const PixItem*, const int16_t*, const uint16_t*,
                                           uint8_t*);
template __global__ void k_pixel<kIdctFloat>(const ImageDesc*, const PixItem*, const int16_t*, const uint16_t*,
                                             uint8_t*);

}  // namespace cjd
