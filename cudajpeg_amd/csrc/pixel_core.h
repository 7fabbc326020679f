// Fused pixel stage of the MI355X baseline-JPEG decoder: dequant + 8x8 iDCT +
// nearest chroma upsample + YCbCr->RGB (or gray copy), one write per output
// byte.  The phases are __host__ __device__ functions over the workgroup's
// LDS state, run by k_pixel (pixel.hip) and lane by lane by the CPU emulator.
//
// Replaces src/pixelTransformGPU.cu:25-142 (iDCT_GPU), :475-511
// (upsampleAndYUVtoRGB_kernel) and :621-677 (iDCT_resample_colourTransform),
// with the edge semantics of CPUonly/src/pixelTransformCPU.c:110-172.
//
// One workgroup takes a run of up to kPixBlocks blocks: consecutive MCUs of
// one MCU row.  Their quantised coefficients (contiguous in scan order) are
// staged in LDS with a per-block descriptor (where the block's 8x8 samples go
// in the component tiles, and which quant table it uses), the blocks are
// inverse-transformed into the component sample tiles, and the tiles are
// colour-converted straight to the interleaved output with aligned 4-byte
// stores.  iDCT variants:
//   kIdctInt   : the reference's integer Chen-Wang iDCT, bit-exact
//                (CPUonly/src/pixelTransformCPU.c:14-107), 8 lanes per block
//   kIdctFloat : separable FP32 iDCT of the dequantised coefficients
//                (T.81 A.3.3), rounded to nearest, 8 lanes per block
//   kIdctMfma  : the same transform as one 64x64 contraction per block on the
//                matrix cores (see the MFMA section below)
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <type_traits>
#if !defined(__HIPCC__)
#include <vector>
#endif

#include "cjd_internal.h"
#include "entropy_core.h"

namespace cjd {

constexpr int kPixThreads = 256;
constexpr int kPixWaves = kPixThreads / 64;

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

// ------------------------------------------------------- FP16 helpers --
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

// FP16 bit pattern of a value that is already exactly representable in FP16.
CJD_HD inline uint16_t half_bits(float h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (_Float16)h);
#else
  uint32_t u;
  memcpy(&u, &h, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const float a = fabsf(h);
  if (a == 0.f) return (uint16_t)sign;
  int e;
  frexpf(a, &e);  // a = m * 2^e
  if (e <= -14) return (uint16_t)(sign | (uint32_t)ldexpf(a, 24));  // subnormal: a / 2^-24
  const uint32_t mant = (uint32_t)ldexpf(a, 11 - e) - 1024u;         // 10 stored bits
  return (uint16_t)(sign | ((uint32_t)(e + 14) << 10) | mant);
#endif
}

// Value of an FP16 bit pattern.
CJD_HD inline float half_value(uint16_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)__builtin_bit_cast(_Float16, b);
#else
  const int e = (b >> 10) & 31, m = b & 1023;
  const float a = e == 0 ? ldexpf((float)m, -24) : ldexpf((float)(m + 1024), e - 25);
  return (b & 0x8000u) ? -a : a;
#endif
}

// Per-block descriptor: offset of the block's top-left sample in `samp`
// (bits 0-13), width of its component tile (bits 14-25), index of its quant
// table among the image's distinct tables (bits 26-27).
CJD_HD inline uint32_t binfo_pack(int base, int tw, int tix) {
  return (uint32_t)base | ((uint32_t)tw << 14) | ((uint32_t)tix << 26);
}
CJD_HD inline int binfo_base(uint32_t v) { return (int)(v & 0x3FFFu); }
CJD_HD inline int binfo_tw(uint32_t v) { return (int)((v >> 14) & 0xFFFu); }
CJD_HD inline int binfo_tix(uint32_t v) { return (int)(v >> 26); }

// LDS of the integer and FP32 forms.  The sample tiles overlay the
// coefficients: the row pass has read every coefficient (behind a barrier)
// before the column pass writes the first sample.
struct PixShared {
  union {
    alignas(16) int16_t coef[kPixBlocks][64];  // quantised coefficients, natural order
    uint8_t samp[kPixBlocks * 64];  // component sample tiles
  };
  union {
    int32_t ti[kPixBlocks][64];  // row-pass output (integer path)
    float tf[kPixBlocks][64];    // row-pass output (float path)
  } t;
  uint16_t q[3][64];              // natural-order quant table per distinct table index
  uint32_t binfo[kPixBlocks];     // per-block descriptor, see binfo_pack
  float cosf[64];                 // kCosF, row-major
};

// LDS of the MFMA form: the coefficients as FP16 bit patterns (the A
// operands), and the sample tiles beside them (waves read A operands and
// write samples concurrently).
struct PixSharedMfma {
  alignas(16) uint16_t coef[kPixBlocks][64];
  alignas(16) uint8_t samp[kPixBlocks * 64];
  uint16_t q[3][64];
  uint32_t binfo[kPixBlocks];
  float cosf[64];
};

// 16-byte piece i (8 coefficients) into LDS: as int16, or as FP16 (exact for
// |c| <= 2048, which holds for the coefficients of any 8-bit-sample image).
typedef uint32_t u32x4 __attribute__((vector_size(16)));
CJD_HD inline void stage_piece(PixShared& sh, int i, const u32x4& v) { __builtin_memcpy(&sh.coef[0][0] + i * 8, &v, 16); }
CJD_HD inline void stage_piece(PixSharedMfma& sh, int i, const u32x4& v) {
  int16_t c[8];
  uint16_t h[8];
  __builtin_memcpy(c, &v, 16);
  for (int k = 0; k < 8; ++k) h[k] = half_bits(round_to_half((float)c[k]));
  __builtin_memcpy(&sh.coef[0][0] + i * 8, h, 16);
}

// Coefficient v at natural index k of block b, and a stream entry (natural
// index << 16 | value; 0xFFFFFFFF = none).
CJD_HD inline void stage_coef(PixShared& sh, int b, uint32_t k, int16_t v) { sh.coef[b][k] = v; }
CJD_HD inline void stage_coef(PixSharedMfma& sh, int b, uint32_t k, int16_t v) {
  sh.coef[b][k] = half_bits(round_to_half((float)v));
}
template <class SH>
CJD_HD inline void stage_entry(SH& sh, int b, uint32_t e) {
  if (e != 0xFFFFFFFFu) stage_coef(sh, b, (e >> 16) & 63u, (int16_t)(e & 0xFFFFu));
}

struct PixCtx {
  const ImageDesc* im;
  PixItem it;
  int nslots, nblk, ncomp, ntab;
  // component c's tile: offset in samp, width (count*sx*8); height sy*8.
  // Scalars for the fixed-component uses; tw(c) / toff(c) for a variable
  // component recompute from the image descriptor, which keeps the context
  // out of private memory.
  int toff0, toff1, toff2, tw0, tw1, tw2;
  int mcu_x0, mcu_y;
  CJD_HD void init(const ImageDesc* images, const PixItem& item) {
    it = item;
    im = images + item.image;
    nslots = im->nslots;
    nblk = (int)it.count * nslots;
    ncomp = im->ncomp;
    ntab = im->ntab;
    tw0 = (int)it.count * im->comp_sx[0] * 8;
    tw1 = ncomp > 1 ? (int)it.count * im->comp_sx[1] * 8 : 0;
    tw2 = ncomp > 2 ? (int)it.count * im->comp_sx[2] * 8 : 0;
    toff0 = 0;
    toff1 = tw0 * im->comp_sy[0] * 8;
    toff2 = toff1 + tw1 * (ncomp > 1 ? im->comp_sy[1] * 8 : 0);
    mcu_y = (int)(it.mcu_first / (uint32_t)im->mcus_x);
    mcu_x0 = (int)(it.mcu_first % (uint32_t)im->mcus_x);
  }
  CJD_HD int tw(int comp) const { return (int)it.count * im->comp_sx[comp] * 8; }
  CJD_HD int toff(int comp) const {
    int off = 0;
    for (int e = 0; e < comp; ++e) off += tw(e) * im->comp_sy[e] * 8;
    return off;
  }
};

// LDS row of block b of the item (b in scan order).  The MFMA form keeps the
// blocks in quant-table order (stable: the item's blocks of lower tables,
// then this table's blocks of earlier MCUs, then its earlier slots in this
// MCU), so that most 16-block tiles of the contraction use one table.
template <class SH>
CJD_HD inline int stage_row(const PixCtx& c, int b) {
  if constexpr (std::is_same<SH, PixSharedMfma>::value) {
    const int m = b / c.nslots, slot = b - m * c.nslots;
    const int t = c.im->comp_tix[c.im->slot_comp[slot]];
    int lower = 0, same = 0, before = 0;
    for (int s2 = 0; s2 < c.nslots; ++s2) {
      const int t2 = c.im->comp_tix[c.im->slot_comp[s2]];
      lower += t2 < t ? 1 : 0;
      same += t2 == t ? 1 : 0;
      before += (t2 == t && s2 < slot) ? 1 : 0;
    }
    return lower * (int)c.it.count + m * same + before;
  } else {
    return b;
  }
}

// The cosine table, once per workgroup.
template <class SH>
CJD_HD inline void pix_stage_cos(SH& sh, int tid) {
  if (tid < 64) sh.cosf[tid] = kCosF[tid >> 3][tid & 7];
}

// Stage one item in two phases with a barrier between them.
//
// pix_stage_zero: every coefficient of the item's blocks to zero, the quant
// tables and the block descriptors.
template <class SH>
CJD_HD inline void pix_stage_zero(SH& sh, int tid, const PixCtx& c, const uint16_t* qtabs) {
  static_assert(kPixBlocks * 8 == 4 * kPixThreads, "four 16-byte pieces per lane");
  const int lim = c.nblk * 8;
  const u32x4 z = {0, 0, 0, 0};
  for (int i = tid; i < lim; i += kPixThreads) __builtin_memcpy(&sh.coef[0][0] + i * 8, &z, 16);
  for (int i = tid; i < c.ntab * 64; i += kPixThreads) {
    const int t = i >> 6;
    sh.q[t][i & 63] = qtabs[(c.im->qtab_base + c.im->tab_dq[t]) * 64 + (i & 63)];
  }
  for (int b = tid; b < c.nblk; b += kPixThreads) {
    const int m = b / c.nslots, slot = b - m * c.nslots;
    const int comp = c.im->slot_comp[slot];
    const int sxf = c.im->comp_sx[comp];
    const int k = slot - c.im->comp_first_slot[comp];
    const int sx = k % sxf, sy = k / sxf;
    const int tw = c.tw(comp);
    sh.binfo[stage_row<SH>(c, b)] = binfo_pack(c.toff(comp) + sy * 8 * tw + (m * sxf + sx) * 8, tw, c.im->comp_tix[comp]);
  }
}

// The entries of 16-byte group v that belong to a block: group entry q is
// the block's entry k0 + q - ph when 0 <= k0 + q - ph < n.
template <class SH>
CJD_HD inline void stage_group(SH& sh, int b, const u32x4& v, uint32_t k0, uint32_t ph, uint32_t n) {
  uint32_t e[4];
  __builtin_memcpy(e, &v, 16);
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q)
    if (k0 + q >= ph && k0 + q < ph + n) stage_entry(sh, b, e[q]);
}

// pix_stage_blocks: lanes 2b and 2b+1 fill block b of the item (blocks are
// the item's MCUs in scan order) from its record, into LDS row stage_row(b).  A block the serial decoder
// wrote is copied from the dense buffer, four 16-byte pieces per lane.  For
// the others the block's AC entries are consecutive entries of a tile stream:
// 16-byte groups of 4, one stream row (4 W words) apart; lane h loads groups
// h, h+2, ... two at a time and stages the entries in them that are the
// block's (an entry's natural index is unique within its block, so the two
// lanes' writes never meet), and the even lane writes the absolute DC value.
// A record whose groups would run past the entry buffer (a corrupt stream)
// stages no AC entries.
template <class SH>
CJD_HD inline void pix_stage_blocks(SH& sh, int tid, const PixCtx& c, const BlockRec* recs, const uint32_t* ebuf,
                                    uint64_t ewords, const int16_t* coefs) {
  const int b = tid >> 1, h = tid & 1;
  if (b >= c.nblk) return;
  const int64_t g = c.im->coef_base + (int64_t)c.it.mcu_first * c.nslots + b;
  const BlockRec r = recs[g];
  const int row = stage_row<SH>(c, b);
  if (r.nw & kRecDense) {
    const int16_t* src = coefs + g * 64 + h * 32;
    u32x4 v0, v1, v2, v3;
    __builtin_memcpy(&v0, src, 16);
    __builtin_memcpy(&v1, src + 8, 16);
    __builtin_memcpy(&v2, src + 16, 16);
    __builtin_memcpy(&v3, src + 24, 16);
    stage_piece(sh, row * 8 + h * 4, v0);
    stage_piece(sh, row * 8 + h * 4 + 1, v1);
    stage_piece(sh, row * 8 + h * 4 + 2, v2);
    stage_piece(sh, row * 8 + h * 4 + 3, v3);
    return;
  }
  const uint32_t n0 = r.nw & 255u, n = n0 < 64 ? n0 : 64, W = 16 * (1 + ((r.nw >> 8) & 3u));
  const uint32_t ph = r.eidx & 3u, ng = n ? (ph + n + 3) >> 2 : 0u;
  const uint64_t g0 = r.eidx & ~3ull, gap = 4ull * W;  // groups are one stream row apart
  if (ng && g0 + (ng - 1) * gap + 4 <= ewords) {
    for (uint32_t g = (uint32_t)h; g < ng; g += 4) {
      u32x4 v0, v1 = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
      __builtin_memcpy(&v0, ebuf + g0 + g * gap, 16);
      if (g + 2 < ng) __builtin_memcpy(&v1, ebuf + g0 + (g + 2) * gap, 16);
      stage_group(sh, row, v0, 4 * g, ph, n);
      stage_group(sh, row, v1, 4 * (g + 2), ph, n);
    }
  }
  if (h == 0) stage_coef(sh, row, 0, r.dc);
}

// Row pass: lane -> (block, row).
template <int kMode>
CJD_HD inline void pix_rows(PixShared& sh, int tid, const PixCtx& c) {
  for (int task = tid; task < c.nblk * 8; task += kPixThreads) {
    const int b = task >> 3, r = task & 7;
    const int16_t* cb = &sh.coef[b][r * 8];
    const uint16_t* qq = &sh.q[binfo_tix(sh.binfo[b])][r * 8];
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
        for (int u = 0; u < 8; ++u) s = fmaf(sh.cosf[u * 8 + x], D[u], s);
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
    const uint32_t info = sh.binfo[b];
    const int tw = binfo_tw(info);
    uint8_t* dst = &sh.samp[binfo_base(info) + x];
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
        for (int v = 0; v < 8; ++v) s = fmaf(sh.cosf[v * 8 + y], D[v], s);
        dst[y * tw] = clip8((int)floorf(s + 0.5f));
      }
    }
  }
}

// Optional planar outputs (reference plane layout, src/format.cu:295-303):
// wave -> row, lane -> column.
template <class SH>
CJD_HD inline void pix_planes(const SH& sh, int tid, const PixCtx& c, uint8_t* out) {
  const int wave = tid >> 6, lane = tid & 63;
  for (int comp = 0; comp < c.ncomp; ++comp) {
    if (c.im->plane_offset[comp] == ~0ull) continue;
    uint8_t* pl = out + c.im->plane_offset[comp];
    const int rows = c.im->comp_sy[comp] * 8;
    const int w = c.tw(comp), toff = c.toff(comp);
    const int px0 = c.mcu_x0 * c.im->comp_sx[comp] * 8, py0 = c.mcu_y * rows;
    for (int yy = wave; yy < rows; yy += kPixWaves)
      for (int xx = lane; xx < w; xx += 64)
        pl[(size_t)(py0 + yy) * c.im->comp_stride[comp] + px0 + xx] = sh.samp[toff + yy * w + xx];
  }
}

// clip8(v >> 8) written as a clamp of v before the shift.  The two forms are
// equal for every int v; this one keeps the device compiler from fusing two
// adjacent clip-and-shift results into one packed 16-bit operation whose
// upper half it then assumes to be zero (on gfx950 it is not: the bits come
// back from the register's previous value).
// clip8 as a 32-bit value.  The colour transform clips ((Y << 8) + t) >> 8
// (CPUonly/src/pixelTransformCPU.c:160-165) as clamp255(Y + (t >> 8)): equal
// for all integers, and no shift follows the clamp (a clamp-then-shift of two
// adjacent results was fused by the compiler into v_ashr_pk_u8_i32, whose
// register was then OR-ed into a word as if its upper half were zero).
CJD_HD inline uint32_t clamp255(int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// 16 bytes at a 4-byte-aligned address (device: one dwordx4 store, which
// needs only dword alignment; host: memcpy).
CJD_HD inline void store_16_a4(uint8_t* p, uint32_t a, uint32_t b, uint32_t d, uint32_t e) {
  const uint32_t v[4] = {a, b, d, e};
  __builtin_memcpy(reinterpret_cast<uint32_t*>(p), v, 16);
}

// Bytes [lo, hi) (0 <= lo < hi <= 16) of the 16 bytes o at a 4-byte-aligned
// address p: the dwords wholly inside as dword stores, the 1-3 bytes at
// either edge as byte and short stores.  At most nine stores, where a byte
// loop takes sixteen: a wave issues every store any of its lanes needs.
CJD_HD inline void store_part16(uint8_t* p, const uint32_t o[4], int lo, int hi) {
  uint32_t* p4 = reinterpret_cast<uint32_t*>(p);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (4 * q >= lo && 4 * q + 4 <= hi) p4[q] = o[q];
  // the dword holding lo, when lo is inside one: bytes lo .. min(hi, its end)
  if (lo & 3) {
    const int q = lo >> 2, e = hi < 4 * q + 4 ? hi : 4 * q + 4;
    const uint32_t v = q == 0 ? o[0] : q == 1 ? o[1] : q == 2 ? o[2] : o[3];
    int k = lo;
    if (k & 1) {
      p[k] = (uint8_t)(v >> (8 * (k & 3)));
      ++k;
    }
    if (k + 2 <= e) {
      *reinterpret_cast<uint16_t*>(p + k) = (uint16_t)(v >> (8 * (k & 3)));
      k += 2;
    }
    if (k < e) p[k] = (uint8_t)(v >> (8 * (k & 3)));
  }
  // the dword holding hi, when hi is inside one the head above did not
  // write: its bytes below hi (lo is at or below its start)
  if ((hi & 3) && !((lo & 3) && (hi >> 2) == (lo >> 2))) {
    const int q = hi >> 2;
    const uint32_t v = q == 0 ? o[0] : q == 1 ? o[1] : q == 2 ? o[2] : o[3];
    int k = 4 * q;
    if (k + 2 <= hi) {
      *reinterpret_cast<uint16_t*>(p + k) = (uint16_t)v;
      k += 2;
    }
    if (k < hi) p[k] = (uint8_t)(v >> (8 * (k & 3)));
  }
}


// Bytes s .. s+3 of the 8-byte value hi:lo (s = 0..3).
CJD_HD inline uint32_t align_bytes(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbyte(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * s));
#endif
}

// Colour conversion / gray copy (CPUonly/src/pixelTransformCPU.c:130-172).
// Each output row segment of the run is written as aligned 16-byte chunks:
// wave -> row, lane -> chunk.  A chunk of interleaved RGB covers bytes of at
// most six pixels, so a lane converts those six and shifts the chunk out of
// their 18 bytes.  The partial chunks at the two ends of a segment (shared
// with neighbouring runs or images) are written byte by byte.
template <class SH>
CJD_HD inline void pix_colour(const SH& sh, int tid, const PixCtx& c, uint8_t* out) {
  const ImageDesc& im = *c.im;
  const int wave = tid >> 6, lane = tid & 63;
  const int x0 = c.mcu_x0 * im.mcu_w, y0 = c.mcu_y * im.mcu_h;
  int x1 = x0 + (int)c.it.count * im.mcu_w;
  if (x1 > im.width) x1 = im.width;
  int y1 = y0 + im.mcu_h;
  if (y1 > im.height) y1 = im.height;
  const int rows = y1 - y0, rw = x1 - x0;
  const int bpp = c.ncomp == 1 ? 1 : 3;
  const int n = rw * bpp;  // bytes per row segment
  const int xs0 = im.comp_xshift[0], ys0 = im.comp_yshift[0];
  const int xs1 = im.comp_xshift[1], ys1 = im.comp_yshift[1];
  const int xs2 = im.comp_xshift[2], ys2 = im.comp_yshift[2];
  for (int yy = wave; yy < rows; yy += kPixWaves) {
    const uint64_t a0 = im.out_offset + ((uint64_t)(y0 + yy) * (uint64_t)im.width + (uint64_t)x0) * (uint64_t)bpp;
    // 16-byte chunks on the row's 4-byte grid: only the first and last
    // chunks are partial, in their first and last dword
    const int mis = (int)(a0 & 3u);
    const int nch = (mis + n + 15) >> 4;
    uint8_t* base = out + (a0 - (uint64_t)mis);
    const uint8_t* srow0 = &sh.samp[c.toff0 + (yy >> ys0) * c.tw0];
    const uint8_t* srow1 = &sh.samp[c.toff1 + (yy >> ys1) * c.tw1];
    const uint8_t* srow2 = &sh.samp[c.toff2 + (yy >> ys2) * c.tw2];
    for (int w = lane; w < nch; w += 64) {
      const int pos = 16 * w - mis;  // segment byte of the chunk's first byte (-3..)
      uint32_t o[4];
      if (bpp == 1) {
        for (int k = 0; k < 4; ++k) {
          uint32_t word = 0;
          for (int j = 0; j < 4; ++j) {
            int q = pos + 4 * k + j;
            q = q < 0 ? 0 : (q >= rw ? rw - 1 : q);
            word |= (uint32_t)srow0[q] << (8 * j);
          }
          o[k] = word;
        }
      } else {
        const int q0 = (pos + 15) / 3 - 5;  // pixel of byte pos (floor, pos >= -15)
        // the six pixels' samples first (their LDS reads in flight together),
        // then (Y << 8) + t >> 8 computed exactly as Y + (t >> 8), with the
        // chroma offsets of 128 folded into the constants:
        //   359 (cr - 128) + 128 = 359 cr - 45824, 454 (cb - 128) + 128 =
        //   454 cb - 57984, 128 - 88 (cb - 128) - 183 (cr - 128) = 34816 -
        //   88 cb - 183 cr
        int sy[6], sb[6], sr[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
          int xx = q0 + e;
          xx = xx < 0 ? 0 : (xx >= rw ? rw - 1 : xx);
          sy[e] = srow0[xx >> xs0];
          sb[e] = srow1[xx >> xs1];
          sr[e] = srow2[xx >> xs2];
        }
        uint32_t px[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) {
          const int r = sy[e] + ((359 * sr[e] - 45824) >> 8);
          const int g = sy[e] + ((34816 - 88 * sb[e] - 183 * sr[e]) >> 8);
          const int b = sy[e] + ((454 * sb[e] - 57984) >> 8);
          px[e] = clamp255(r) | (clamp255(g) << 8) | (clamp255(b) << 16);
        }
        // the six pixels' 18 bytes as dwords, then the 16 from byte pos - 3 q0
        const uint32_t d0 = px[0] | (px[1] << 24), d1 = (px[1] >> 8) | (px[2] << 16);
        const uint32_t d2 = (px[2] >> 16) | (px[3] << 8), d3 = px[4] | (px[5] << 24), d4 = px[5] >> 8;
        const uint32_t sft = (uint32_t)(pos - 3 * q0);
        o[0] = align_bytes(d1, d0, sft);
        o[1] = align_bytes(d2, d1, sft);
        o[2] = align_bytes(d3, d2, sft);
        o[3] = align_bytes(d4, d3, sft);
      }
      uint8_t* dst = base + 16 * w;
      if (pos >= 0 && pos + 16 <= n)
        store_16_a4(dst, o[0], o[1], o[2], o[3]);
      else
        store_part16(dst, o, pos < 0 ? -pos : 0, n - pos < 16 ? n - pos : 16);
    }
  }
}

// ----------------------------------------------------------- MFMA iDCT --
// The inverse DCT of a block as one 64x64 contraction,
//   s[p] = 128 + sum_k c[k] * M_t[k][p],   M_t[k][p] = q_t[k] * Cu[u][x] * Cv[v][y]
// (k = 8v + u natural coefficient index, p = 8y + x pixel index, q_t the
// quant table), evaluated on the matrix cores with v_mfma_f32_16x16x32_f16:
// the quantised coefficients (|c| <= 2047) are exact in FP16, M_t is split
// into FP16 hi + lo parts and the products accumulate in FP32.  A tile is 16
// blocks x 64 pixels; the image's distinct quant tables are stacked along K
// (a block's coefficients are zero in the K range of the other tables), so a
// tile with mixed luma/chroma blocks is one pass.
//
// Lane layout of v_mfma_f32_16x16x32_f16 (checked on the device): A[row][k]
// lane l holds row l%16, k = 8*(l/16) + i; B[k][col] lane l holds col l%16,
// the same k; C lane l holds col l%16, rows 4*(l/16) + i.  The contraction
// runs as C[pixel][block] = M_t^T x coef^T: with K half h (coefficients
// 32h..32h+31) and N tile n (pixels 16n..16n+15), lane l's A element i is
// M_t[32h + 8(l/16) + i][16n + l%16] and its B element i is coefficient
// 32h + 8(l/16) + i of block l%16; its 4 results are pixels 16n + 4(l/16) + i
// of that block, 4 consecutive samples of one row.

// Table operand (A) of lane l for quant table q (natural order), N tile n and
// K half h: its 8 elements of M = (q[k] * Cu[u][x]) * Cv[v][y] in FP32, split
// into FP16 hi and lo bit patterns.
CJD_HD inline void mfma_b_lane(const uint16_t* q, const float* cosf, int n, int h, int l, uint16_t hb[8],
                               uint16_t lb[8]) {
  const int p = n * 16 + (l & 15), x = p & 7, y = p >> 3;
  const int kb = h * 32 + 8 * (l >> 4);  // k = kb + i: v = kb / 8, u = i
  const float cv = cosf[(kb >> 3) * 8 + y];
  for (int i = 0; i < 8; ++i) {
    const float m = (float)q[kb + i] * cosf[i * 8 + x] * cv;
    const float hi = round_to_half(m);
    hb[i] = half_bits(hi);
    lb[i] = half_bits(round_to_half(m - hi));
  }
}

// Host emulation of the device MFMA phase, all table passes: the same
// operands (the table operand from mfma_b_lane, the block operand from the
// FP16 coefficients in LDS, blocks in table order) held by the same lanes,
// the same tiles skipped and the same masking, with each 16x16x32 product
// summed in k order in FP32 (the device's summation order inside one MFMA is
// the hardware's), and each lane's 4 results written as one sample word.
// Checks the operand layout, the table order and the output addressing on
// the CPU.
#if !defined(__HIPCC__)
inline void pix_idct_mfma_emul(PixSharedMfma& sh, const PixCtx& c) {
  const int ntiles = (c.nblk + 15) >> 4;
  const int last = c.nblk - 1;
  for (int t0 = 0; t0 < c.ntab; t0 += 2) {
    const int nt = c.ntab - t0 < 2 ? c.ntab - t0 : 2;
    for (int n = 0; n < 4; ++n) {  // the wave that owns N tile n
      float A[2][2][2][64][8];     // [table][h][hi, lo][lane][element]
      for (int tl = 0; tl < nt; ++tl)
        for (int h = 0; h < 2; ++h)
          for (int l = 0; l < 64; ++l) {
            uint16_t hb[8], lb[8];
            mfma_b_lane(sh.q[t0 + tl], sh.cosf, n, h, l, hb, lb);
            for (int i = 0; i < 8; ++i) {
              A[tl][h][0][l][i] = half_value(hb[i]);
              A[tl][h][1][l][i] = half_value(lb[i]);
            }
          }
      for (int tile = 0; tile < ntiles; ++tile) {
        const int f = tile * 16;
        // the pair's tables (the device takes tiles two at a time): of the
        // first and last block of the pair, in table order
        const int fp = f & ~31, ep = fp + 31 <= last ? fp + 31 : last;
        const int tlo = binfo_tix(sh.binfo[fp]) - t0, thi = binfo_tix(sh.binfo[ep]) - t0;
        float bv[64][2][8];
        int tix[64], blk[64];
        for (int l = 0; l < 64; ++l) {
          const int pj = f + (l & 15);
          blk[l] = pj <= last ? pj : last;
          tix[l] = pj <= last ? binfo_tix(sh.binfo[blk[l]]) : -1;
          for (int h = 0; h < 2; ++h)
            for (int i = 0; i < 8; ++i) bv[l][h][i] = half_value(sh.coef[blk[l]][h * 32 + 8 * (l >> 4) + i]);
        }
        float acc[64][4] = {};
        for (int tl = 0; tl < nt; ++tl) {
          // one table in the pair: only its pass, unmasked; else every pass, masked
          const bool single = tlo == thi;
          if (single && tl != tlo) continue;
          for (int h = 0; h < 2; ++h)
            for (int hl = 0; hl < 2; ++hl) {
              float prod[16][16];  // [pixel row][block col]
              for (int row = 0; row < 16; ++row)
                for (int col = 0; col < 16; ++col) {
                  float s = 0.f;
                  for (int k = 0; k < 32; ++k) {
                    const int la = row + 16 * (k >> 3), lb = col + 16 * (k >> 3);
                    const float b = (single || tix[lb] == t0 + tl) ? bv[lb][h][k & 7] : 0.f;
                    s += A[tl][h][hl][la][k & 7] * b;
                  }
                  prod[row][col] = s;
                }
              for (int l = 0; l < 64; ++l)
                for (int i = 0; i < 4; ++i) acc[l][i] += prod[4 * (l >> 4) + i][l & 15];
            }
        }
        for (int l = 0; l < 64; ++l) {
          if (tix[l] < t0 || tix[l] >= t0 + nt) continue;
          const uint32_t info = sh.binfo[blk[l]];
          const int yo = 2 * n + (l >> 5), xo = 4 * ((l >> 4) & 1);
          uint8_t* dst = &sh.samp[binfo_base(info) + yo * binfo_tw(info) + xo];
          if ((binfo_base(info) + yo * binfo_tw(info) + xo) & 3) __builtin_trap();  // the device writes a word
          for (int i = 0; i < 4; ++i) dst[i] = clip8((int)floorf(acc[l][i] + 128.5f));
        }
      }
    }
  }
}
#endif

}  // namespace cjd
