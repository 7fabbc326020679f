// Fused pixel stage of the MI355X baseline-JPEG decoder: dequant + 8x8 iDCT +
// nearest chroma upsample + YCbCr->RGB (or gray copy), one write per output
// pixel.  The phases are __host__ __device__ functions over the workgroup's
// LDS state, run by k_pixel (pixel.hip) and lane by lane by the CPU emulator.
//
// Replaces src/pixelTransformGPU.cu:25-142 (iDCT_GPU), :475-511
// (upsampleAndYUVtoRGB_kernel) and :621-677 (iDCT_resample_colourTransform),
// with the edge semantics of CPUonly/src/pixelTransformCPU.c:110-172.
//
// One workgroup takes a run of consecutive MCUs of one MCU row.  Their
// quantised coefficients (contiguous in scan order) are staged in LDS, each
// block is inverse-transformed by 8 lanes (rows, then columns) into per-
// component sample tiles, and the tiles are colour-converted straight to the
// interleaved output.  iDCT variants:
//   kIdctInt   : the reference's integer Chen-Wang iDCT, bit-exact
//                (CPUonly/src/pixelTransformCPU.c:14-107)
//   kIdctFloat : separable FP32 iDCT of the dequantised coefficients
//                (T.81 A.3.3), rounded to nearest
#pragma once
#include <math.h>
#include <stdint.h>

#include "cjd_internal.h"
#include "entropy_core.h"

namespace cjd {

constexpr int kPixThreads = 256;

// CPUonly/include/pixelTransformCPU.h:5-10
#define W1 2841
#define W2 2676
#define W3 2408
#define W5 1609
#define W6 1108
#define W7 565

CJD_HD inline uint8_t clip8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

// Row pass, CPUonly/src/pixelTransformCPU.c:14-57 without the all-zero
// shortcut (which yields identical values).
CJD_HD inline void idct_row_int(const int* D, int* O) {
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
CJD_HD inline void idct_col_int(const int* D, uint8_t* out) {
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
constexpr float kCosF[8][8] = {
    {0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f, 0.35355339f},
    {0.49039264f, 0.41573481f, 0.27778512f, 0.09754516f, -0.09754516f, -0.27778512f, -0.41573481f, -0.49039264f},
    {0.46193977f, 0.19134172f, -0.19134172f, -0.46193977f, -0.46193977f, -0.19134172f, 0.19134172f, 0.46193977f},
    {0.41573481f, -0.09754516f, -0.49039264f, -0.27778512f, 0.27778512f, 0.49039264f, 0.09754516f, -0.41573481f},
    {0.35355339f, -0.35355339f, -0.35355339f, 0.35355339f, 0.35355339f, -0.35355339f, -0.35355339f, 0.35355339f},
    {0.27778512f, -0.49039264f, 0.09754516f, 0.41573481f, -0.41573481f, -0.09754516f, 0.49039264f, -0.27778512f},
    {0.19134172f, -0.46193977f, 0.46193977f, -0.19134172f, -0.19134172f, 0.46193977f, -0.46193977f, 0.19134172f},
    {0.09754516f, -0.27778512f, 0.41573481f, -0.49039264f, 0.49039264f, -0.41573481f, 0.27778512f, -0.09754516f}};


struct PixShared {
  int16_t coef[kPixBlocks][64];  // quantised coefficients, natural order
  union {
    int32_t ti[kPixBlocks][64];  // row-pass output (integer path)
    float tf[kPixBlocks][64];    // row-pass output (float path)
  } t;
  uint8_t samp[kPixBlocks * 64];  // component sample tiles
  uint16_t q[3][64];              // natural-order quant table per component
};

struct PixCtx {
  const ImageDesc* im;
  PixItem it;
  int nslots, nblk, ncomp;
  int toff[3], tw[3];  // component c's tile: offset in samp, width (count*sx*8); height sy*8
  int mcu_x0, mcu_y;
  CJD_HD void init(const ImageDesc* images, const PixItem& item) {
    it = item;
    im = images + item.image;
    nslots = im->nslots;
    nblk = (int)it.count * nslots;
    ncomp = im->ncomp;
    int off = 0;
    for (int c = 0; c < 3; ++c) {
      toff[c] = off;
      tw[c] = c < ncomp ? (int)it.count * im->comp_sx[c] * 8 : 0;
      if (c < ncomp) off += tw[c] * im->comp_sy[c] * 8;
    }
    mcu_y = (int)(it.mcu_first / (uint32_t)im->mcus_x);
    mcu_x0 = (int)(it.mcu_first % (uint32_t)im->mcus_x);
  }
};

// Stage coefficients (contiguous: MCUs in scan order) and quant tables.
CJD_HD inline void pix_stage(PixShared& sh, int tid, const PixCtx& c, const int16_t* coefs, const uint16_t* qtabs) {
  const int64_t first = c.im->coef_base + (int64_t)c.it.mcu_first * c.nslots;
  // 16-byte pieces, copied with memcpy (no type punning of the int16 data)
  const int16_t* src = coefs + first * 64;
  int16_t* dst = &sh.coef[0][0];
  for (int i = tid; i < c.nblk * 8; i += kPixThreads) __builtin_memcpy(dst + i * 8, src + i * 8, 16);
  for (int i = tid; i < c.ncomp * 64; i += kPixThreads) {
    const int comp = i >> 6;
    sh.q[comp][i & 63] = qtabs[(c.im->qtab_base + c.im->comp_dq[comp]) * 64 + (i & 63)];
  }
}

// Row pass: lane -> (block, row).
template <int kMode>
CJD_HD inline void pix_rows(PixShared& sh, int tid, const PixCtx& c) {
  for (int task = tid; task < c.nblk * 8; task += kPixThreads) {
    const int b = task >> 3, r = task & 7;
    const int comp = c.im->slot_comp[b % c.nslots];
    const int16_t* cb = &sh.coef[b][r * 8];
    const uint16_t* qq = &sh.q[comp][r * 8];
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
}

// Column pass: lane -> (block, column); samples into the component tile.
template <int kMode>
CJD_HD inline void pix_cols(PixShared& sh, int tid, const PixCtx& c) {
  for (int task = tid; task < c.nblk * 8; task += kPixThreads) {
    const int b = task >> 3, x = task & 7;
    const int m = b / c.nslots, slot = b - m * c.nslots;
    const int comp = c.im->slot_comp[slot];
    const int k = slot - c.im->comp_first_slot[comp];
    const int sx = k % c.im->comp_sx[comp], sy = k / c.im->comp_sx[comp];
    const int tw = c.tw[comp];
    uint8_t* dst = &sh.samp[c.toff[comp] + (sy * 8) * tw + (m * c.im->comp_sx[comp] + sx) * 8 + x];
    if (kMode == kIdctInt) {
      int D[8];
      uint8_t o[8];
      for (int v = 0; v < 8; ++v) D[v] = sh.t.ti[b][v * 8 + x];
      idct_col_int(D, o);
      for (int y = 0; y < 8; ++y) dst[y * tw] = o[y];
    } else {
      float D[8];
      for (int v = 0; v < 8; ++v) D[v] = sh.t.tf[b][v * 8 + x];
      for (int y = 0; y < 8; ++y) {
        float s = 128.f;
        for (int v = 0; v < 8; ++v) s = fmaf(kCosF[v][y], D[v], s);
        dst[y * tw] = clip8((int)floorf(s + 0.5f));
      }
    }
  }
}

// Optional planar outputs (reference plane layout, src/format.cu:295-303).
CJD_HD inline void pix_planes(const PixShared& sh, int tid, const PixCtx& c, uint8_t* out) {
  for (int comp = 0; comp < c.ncomp; ++comp) {
    if (c.im->plane_offset[comp] == ~0ull) continue;
    uint8_t* pl = out + c.im->plane_offset[comp];
    const int rows = c.im->comp_sy[comp] * 8, w = c.tw[comp];
    const int px0 = c.mcu_x0 * c.im->comp_sx[comp] * 8, py0 = c.mcu_y * rows;
    for (int i = tid; i < rows * w; i += kPixThreads) {
      const int yy = i / w, xx = i - yy * w;
      pl[(size_t)(py0 + yy) * c.im->comp_stride[comp] + px0 + xx] = sh.samp[c.toff[comp] + yy * w + xx];
    }
  }
}

// Colour conversion / gray copy, one output byte per lane per step
// (CPUonly/src/pixelTransformCPU.c:130-172).
CJD_HD inline void pix_colour(const PixShared& sh, int tid, const PixCtx& c, uint8_t* out) {
  const ImageDesc& im = *c.im;
  const int x0 = c.mcu_x0 * im.mcu_w, y0 = c.mcu_y * im.mcu_h;
  int x1 = x0 + (int)c.it.count * im.mcu_w;
  if (x1 > im.width) x1 = im.width;
  int y1 = y0 + im.mcu_h;
  if (y1 > im.height) y1 = im.height;
  const int rw = x1 - x0, rows = y1 - y0;
  uint8_t* o = out + im.out_offset;
  if (c.ncomp == 1) {
    for (int i = tid; i < rw * rows; i += kPixThreads) {
      const int yy = i / rw, xx = i - yy * rw;
      o[(size_t)(y0 + yy) * im.width + x0 + xx] = sh.samp[yy * c.tw[0] + xx];
    }
    return;
  }
  const int xs0 = im.comp_xshift[0], ys0 = im.comp_yshift[0];
  const int xs1 = im.comp_xshift[1], ys1 = im.comp_yshift[1];
  const int xs2 = im.comp_xshift[2], ys2 = im.comp_yshift[2];
  const int nbytes = rw * 3;
  for (int i = tid; i < nbytes * rows; i += kPixThreads) {
    const int yy = i / nbytes, bx = i - yy * nbytes;
    const int xx = bx / 3, ch = bx - xx * 3;
    const int Y = (int)sh.samp[c.toff[0] + (yy >> ys0) * c.tw[0] + (xx >> xs0)] << 8;
    const int cb = (int)sh.samp[c.toff[1] + (yy >> ys1) * c.tw[1] + (xx >> xs1)] - 128;
    const int cr = (int)sh.samp[c.toff[2] + (yy >> ys2) * c.tw[2] + (xx >> xs2)] - 128;
    int v;
    if (ch == 0) v = (Y + 359 * cr + 128) >> 8;
    else if (ch == 1) v = (Y - 88 * cb - 183 * cr + 128) >> 8;
    else v = (Y + 454 * cb + 128) >> 8;
    o[((size_t)(y0 + yy) * im.width + x0) * 3 + bx] = clip8(v);
  }
}

}  // namespace cjd

namespace cjd {

// ----------------------------------------------------------- MFMA iDCT --
// The inverse DCT of a block as one 64x64 contraction,
//   s[p] = 128 + sum_k c[k] * M_t[k][p],   M_t[k][p] = q_t[k] * Cu[u][x] * Cv[v][y]
// (k = 8v + u natural coefficient index, p = 8y + x pixel index, q_t the
// quant table), evaluated on the matrix cores: the quantised coefficients
// (|c| <= 2047) are exact in FP16, M_t is split into FP16 hi + lo parts and
// the products accumulate in FP32.  Tiles of 16 blocks x 64 pixels; one pass
// per distinct quant table of the image, rows of other tables masked to zero.
//
// Lane layout of v_mfma_f32_16x16x32_f16 (checked on the device): A[row][k]
// lane l holds row l%16, k = 8*(l/16) + i; B[k][col] lane l holds col l%16,
// the same k; C lane l holds col l%16, rows 4*(l/16) + i.

// Round a float to the nearest FP16 value (ties to even), as a float.
CJD_HD inline float round_to_half(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)(_Float16)x;
#else
  if (x == 0.f) return 0.f;
  int e;
  frexpf(x, &e);                   // x = m * 2^e, 0.5 <= |m| < 1
  int qe = e - 11;                 // quantum of an 11-bit significand
  if (qe < -24) qe = -24;          // subnormal quantum
  return ldexpf(nearbyintf(ldexpf(x, -qe)), qe);
#endif
}

// M_t split for lane l of the B operand: hi/lo[kk][n][i] for K-step kk,
// N-tile n, element i.
CJD_HD inline void mfma_basis(const uint16_t* q, int l, float hi[2][4][8], float lo[2][4][8]) {
  const int x = l & 7, ysub = (l & 15) >> 3, vsub = l >> 4;
  for (int kk = 0; kk < 2; ++kk)
    for (int n = 0; n < 4; ++n) {
      const float cv = kCosF[kk * 4 + vsub][2 * n + ysub];
      for (int i = 0; i < 8; ++i) {
        const float m = (float)q[kk * 32 + vsub * 8 + i] * kCosF[i][x] * cv;
        const float h = round_to_half(m);
        hi[kk][n][i] = h;
        lo[kk][n][i] = round_to_half(m - h);
      }
    }
}

// Write one contraction result (pixel p of block b) into the sample tile.
CJD_HD inline void mfma_store(PixShared& sh, const PixCtx& c, int b, int p, float v) {
  const int m = b / c.nslots, slot = b - m * c.nslots;
  const int comp = c.im->slot_comp[slot];
  const int k = slot - c.im->comp_first_slot[comp];
  const int sx = k % c.im->comp_sx[comp], sy = k / c.im->comp_sx[comp];
  const int tw = c.tw[comp];
  const int y = p >> 3, x = p & 7;
  sh.samp[c.toff[comp] + (sy * 8 + y) * tw + (m * c.im->comp_sx[comp] + sx) * 8 + x] =
      clip8((int)floorf(v + 128.5f));
}

// Host restatement of the same contraction for the CPU emulator: identical
// hi/lo split and FP32 products, accumulated in a fixed order.
inline void pix_idct_hilo_host(PixShared& sh, const PixCtx& c) {
  for (int b = 0; b < c.nblk; ++b) {
    const int comp = c.im->slot_comp[b % c.nslots];
    const uint16_t* q = sh.q[comp];
    for (int p = 0; p < 64; ++p) {
      const int y = p >> 3, x = p & 7;
      float s = 0.f;
      for (int k = 0; k < 64; ++k) {
        const float m = (float)q[k] * kCosF[k & 7][x] * kCosF[k >> 3][y];
        const float h = round_to_half(m);
        const float lo = round_to_half(m - h);
        s += (float)sh.coef[b][k] * h;
        s += (float)sh.coef[b][k] * lo;
      }
      mfma_store(sh, c, b, p, s);
    }
  }
}

}  // namespace cjd
