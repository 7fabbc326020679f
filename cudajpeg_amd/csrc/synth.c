/*
 * Seeded synthetic-image generator and baseline JPEG encoder.
 *
 * This produces the decoder's input files and, beside each file, the
 * encoder-side ground truth: the quantised coefficients of every 8x8 block
 * (natural row-major order, blocks in scan order) and the bit offset at which
 * every block starts in the unstuffed entropy-coded segment.
 *
 * Content per channel: 128 + sum of 6 random 2-D sinusoids (period 16-256 px,
 * random orientation and phase, amplitudes a Dirichlet split of `amp_total`)
 * plus Gaussian noise of std `sigma`, optionally with a fraction of 64x64 tiles
 * replaced by uniform random bytes, clamped to uint8.
 *
 * Encoding: JFIF baseline, 8-bit, 1 or 3 components, luma sampling h x v with
 * chroma 1x1, IJG-scaled Annex K quantisation tables, Annex K Huffman tables,
 * one interleaved scan, optional restart interval.  Everything is
 * deterministic in `seed`.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth.h"

/* ---------------------------------------------------------------- PRNG -- */
/* splitmix64 to seed, xoshiro256** to draw (public-domain algorithms). */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
typedef struct { uint64_t s[4]; } rng_t;
static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
  uint64_t x = seed * 0x2545F4914F6CDD1DULL + stream * 0x9E3779B97F4A7C15ULL + 1;
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return res;
}
static inline double rng_unif(rng_t *r) { return (rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

/* Inverse normal CDF (Acklam's rational approximation, |rel err| < 1.2e-9),
 * used once to tabulate 65536 equiprobable Gaussian quantiles. */
static double inv_norm(double p) {
  static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
  static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
  static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                             3.754408661907416e+00};
  double q, r;
  if (p < 0.02425) {
    q = sqrt(-2 * log(p));
    return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  }
  if (p > 1 - 0.02425) {
    q = sqrt(-2 * log(1 - p));
    return -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
           ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  }
  q = p - 0.5; r = q * q;
  return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
         (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
}

static float g_norm_tab[65536];
static volatile int g_norm_ready = 0;
static void norm_init(void) {
  if (g_norm_ready) return;
  for (int i = 0; i < 65536; i++) g_norm_tab[i] = (float)inv_norm((i + 0.5) / 65536.0);
  g_norm_ready = 1;
}

/* ------------------------------------------------------------- content -- */
#define NSIN 6
int synth_pixels(const synth_params *p, uint8_t *out) {
  const int W = p->width, H = p->height, C = p->ncomp == 1 ? 1 : 3;
  if (W <= 0 || H <= 0) return -1;
  norm_init();
  rng_t r;
  rng_seed(&r, p->seed, 0xC0FFEE);
  double ax[3][NSIN], by[3][NSIN], ph[3][NSIN], amp[3][NSIN];
  for (int c = 0; c < C; c++) {
    double e[NSIN], es = 0;
    for (int i = 0; i < NSIN; i++) {
      double P = 16.0 + 240.0 * rng_unif(&r);
      double th = M_PI * rng_unif(&r);
      ph[c][i] = 2 * M_PI * rng_unif(&r);
      ax[c][i] = 2 * M_PI * cos(th) / P;
      by[c][i] = 2 * M_PI * sin(th) / P;
      e[i] = -log(1.0 - rng_unif(&r)); /* Dirichlet(1..1) via exponentials */
      es += e[i];
    }
    for (int i = 0; i < NSIN; i++) amp[c][i] = p->amp_total * e[i] / es;
  }
  /* sin(a x + b y + phi) = sin(a x) cos(b y + phi) + cos(a x) sin(b y + phi) */
  float *sx = (float *)malloc(sizeof(float) * (size_t)W * NSIN * 2);
  float *rowv = (float *)malloc(sizeof(float) * (size_t)W);
  if (!sx || !rowv) { free(sx); free(rowv); return -2; }
  /* random 64x64 tiles: one uniform draw per tile decides */
  const int TW = (W + 63) / 64, TH = (H + 63) / 64;
  uint8_t *tile = (uint8_t *)calloc((size_t)TW * TH, 1);
  if (!tile) { free(sx); free(rowv); return -2; }
  if (p->random_tile_frac > 0) {
    rng_t tr;
    rng_seed(&tr, p->seed, 0x711E5);
    for (int t = 0; t < TW * TH; t++) tile[t] = rng_unif(&tr) < p->random_tile_frac;
  }
  const float sig = (float)p->sigma;
  for (int c = 0; c < C; c++) {
    for (int i = 0; i < NSIN; i++)
      for (int x = 0; x < W; x++) {
        sx[(size_t)(2 * i) * W + x] = (float)(amp[c][i] * sin(ax[c][i] * x));
        sx[(size_t)(2 * i + 1) * W + x] = (float)(amp[c][i] * cos(ax[c][i] * x));
      }
    for (int y = 0; y < H; y++) {
      float cy[NSIN], sy[NSIN];
      for (int i = 0; i < NSIN; i++) {
        cy[i] = (float)cos(by[c][i] * y + ph[c][i]);
        sy[i] = (float)sin(by[c][i] * y + ph[c][i]);
      }
      for (int x = 0; x < W; x++) {
        float v = 128.0f;
        for (int i = 0; i < NSIN; i++)
          v += sx[(size_t)(2 * i) * W + x] * cy[i] + sx[(size_t)(2 * i + 1) * W + x] * sy[i];
        rowv[x] = v;
      }
      rng_t nr;
      rng_seed(&nr, p->seed ^ ((uint64_t)(c + 1) << 56), (uint64_t)y + 1);
      uint64_t bits = 0;
      int nb = 0;
      uint8_t *o = out + (size_t)y * W * C + c;
      for (int x = 0; x < W; x++) {
        if (nb == 0) { bits = rng_next(&nr); nb = 4; }
        float v = rowv[x] + sig * g_norm_tab[bits & 0xFFFF];
        bits >>= 16; nb--;
        if (tile[(y >> 6) * TW + (x >> 6)]) {
          v = (float)(rng_next(&nr) >> 56);
        }
        int iv = (int)lrintf(v);
        o[(size_t)x * C] = (uint8_t)(iv < 0 ? 0 : iv > 255 ? 255 : iv);
      }
    }
  }
  free(sx); free(rowv); free(tile);
  return 0;
}

/* ------------------------------------------------------ Annex K tables -- */
static const uint8_t k_lum_q[64] = {
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55,
    14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51, 87, 80, 62,
    18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
static const uint8_t k_chr_q[64] = {
    17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
    24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
/* zigzag index -> natural index */
static const uint8_t k_zz[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15,
    23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

static const uint8_t k_dc_lum_bits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t k_dc_chr_bits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
static const uint8_t k_dc_vals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
static const uint8_t k_ac_lum_bits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
static const uint8_t k_ac_lum_vals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61,
    0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52,
    0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25,
    0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45,
    0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,
    0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6,
    0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3,
    0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8,
    0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
static const uint8_t k_ac_chr_bits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
static const uint8_t k_ac_chr_vals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61,
    0x71, 0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33,
    0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18,
    0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44,
    0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63,
    0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97,
    0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4,
    0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
    0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7,
    0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

void synth_quant_table16(int quality, int chroma, int maxv, uint16_t out_natural[64]) {
  if (quality < 1) quality = 1;
  if (quality > 100) quality = 100;
  int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  const uint8_t *base = chroma ? k_chr_q : k_lum_q;
  for (int i = 0; i < 64; i++) {
    long v = ((long)base[i] * scale + 50) / 100;
    out_natural[i] = (uint16_t)(v < 1 ? 1 : v > maxv ? maxv : v);
  }
}

void synth_quant_table(int quality, int chroma, uint8_t out_natural[64]) {
  uint16_t t[64];
  synth_quant_table16(quality, chroma, 255, t);
  for (int i = 0; i < 64; i++) out_natural[i] = (uint8_t)t[i];
}

/* Encoder-side code tables (T.81 C.2 canonical construction). */
/* "Long codes" AC code lengths (huff = 1): the Annex K symbol order, 10
 * short codes, 60 codes of 11 bits and the rest of 16 bits -- a valid table
 * whose long codes need more second-level lookup tables than common tables. */
static const uint8_t k_ac_long_bits[16] = {0, 2, 2, 2, 2, 2, 0, 0, 0, 0, 60, 0, 0, 0, 0, 92};
static const uint8_t k_ac_long_bits_chr[16] = {0, 2, 2, 2, 2, 2, 0, 0, 0, 0, 60, 0, 0, 0, 0, 92};

typedef struct { uint16_t code[256]; uint8_t len[256]; } ehuff_t;
static void build_ehuff(const uint8_t bits[16], const uint8_t *vals, ehuff_t *t) {
  memset(t, 0, sizeof(*t));
  int k = 0, code = 0;
  for (int l = 1; l <= 16; l++) {
    for (int i = 0; i < bits[l - 1]; i++, k++) {
      t->code[vals[k]] = (uint16_t)code++;
      t->len[vals[k]] = (uint8_t)l;
    }
    code <<= 1;
  }
}

/* ----------------------------------------------------------- bit writer -- */
typedef struct {
  uint8_t *out; long cap, n;
  uint64_t acc; int nacc;   /* pending bits, right-aligned */
  uint64_t raw_bits;        /* bits written to the unstuffed stream */
  int overflow;
} bw_t;
static inline void bw_byte(bw_t *w, uint8_t b) {
  if (w->n + 2 > w->cap) { w->overflow = 1; return; }
  w->out[w->n++] = b;
  if (b == 0xFF) w->out[w->n++] = 0x00;
}
static inline void bw_put(bw_t *w, uint32_t bits, int n) {
  if (!n) return;
  w->acc = (w->acc << n) | (bits & ((1u << n) - 1));
  w->nacc += n;
  w->raw_bits += n;
  while (w->nacc >= 8) {
    w->nacc -= 8;
    bw_byte(w, (uint8_t)(w->acc >> w->nacc));
  }
}
static void bw_flush_ones(bw_t *w) {
  if (w->nacc) {
    int pad = 8 - w->nacc;
    w->acc = (w->acc << pad) | ((1u << pad) - 1);
    w->nacc = 0;
    w->raw_bits += pad;
    bw_byte(w, (uint8_t)w->acc);
  }
}
static void raw(bw_t *w, uint8_t b) {
  if (w->n + 1 > w->cap) { w->overflow = 1; return; }
  w->out[w->n++] = b;
}
static void raw16(bw_t *w, int v) { raw(w, (uint8_t)(v >> 8)); raw(w, (uint8_t)v); }

static inline int nbits_of(int v) {
  int a = v < 0 ? -v : v, n = 0;
  while (a) { n++; a >>= 1; }
  return n;
}

/* ------------------------------------------------------------ encoder -- */
/* Exact 8x8 forward DCT (T.81 A.3.3) on level-shifted samples. */
static double g_cos[8][8];
static volatile int g_cos_ready = 0;
static void cos_init(void) {
  if (g_cos_ready) return;
  for (int u = 0; u < 8; u++)
    for (int x = 0; x < 8; x++)
      g_cos[u][x] = (u == 0 ? sqrt(0.125) : 0.5) * cos((2 * x + 1) * u * M_PI / 16.0);
  g_cos_ready = 1;
}
static void fdct_quant(const float blk[64], const uint16_t qnat[64], int16_t out[64]) {
  double tmp[64];
  for (int y = 0; y < 8; y++)
    for (int u = 0; u < 8; u++) {
      double s = 0;
      for (int x = 0; x < 8; x++) s += g_cos[u][x] * blk[y * 8 + x];
      tmp[y * 8 + u] = s;
    }
  for (int v = 0; v < 8; v++)
    for (int u = 0; u < 8; u++) {
      double s = 0;
      for (int y = 0; y < 8; y++) s += g_cos[v][y] * tmp[y * 8 + u];
      long q = lround(s / qnat[v * 8 + u]);
      long lim = (v == 0 && u == 0) ? 2047 : 1023;
      if (q > lim) q = lim;
      if (q < -lim) q = -lim;
      out[v * 8 + u] = (int16_t)q;
    }
}

/* quant table of component c: Y 0, Cb 1, Cr 1 (or 2 with three tables) */
static int comp_qtab(const synth_params *p, int c) { return c == 0 ? 0 : (c == 2 && p->qtables == 3 ? 2 : 1); }

static void put_header(bw_t *w, const synth_params *p, uint16_t q[3][64]) {
  const int C = p->ncomp == 1 ? 1 : 3;
  raw(w, 0xFF); raw(w, 0xD8);
  /* APP0 JFIF */
  raw(w, 0xFF); raw(w, 0xE0); raw16(w, 16);
  raw(w, 'J'); raw(w, 'F'); raw(w, 'I'); raw(w, 'F'); raw(w, 0);
  raw(w, 1); raw(w, 1); raw(w, 0); raw16(w, 1); raw16(w, 1); raw(w, 0); raw(w, 0);
  /* DQT, zigzag order */
  for (int t = 0; t < (C == 1 ? 1 : (p->qtables == 3 ? 3 : 2)); t++) {
    if (p->dqt16) {  /* Pq = 1: 16-bit entries, big-endian */
      raw(w, 0xFF); raw(w, 0xDB); raw16(w, 131); raw(w, (uint8_t)(0x10 | t));
      for (int k = 0; k < 64; k++) raw16(w, q[t][k_zz[k]]);
    } else {
      raw(w, 0xFF); raw(w, 0xDB); raw16(w, 67); raw(w, (uint8_t)t);
      for (int k = 0; k < 64; k++) raw(w, (uint8_t)q[t][k_zz[k]]);
    }
  }
  /* SOF0 */
  raw(w, 0xFF); raw(w, 0xC0); raw16(w, 8 + 3 * C); raw(w, 8);
  raw16(w, p->height); raw16(w, p->width); raw(w, (uint8_t)C);
  for (int c = 0; c < C; c++) {
    raw(w, (uint8_t)(c + 1));
    if (c == 0) raw(w, (uint8_t)((C == 1 ? 1 : p->h_samp) << 4 | (C == 1 ? 1 : p->v_samp)));
    else raw(w, 0x11);
    raw(w, (uint8_t)comp_qtab(p, c));
  }
  /* DHT */
  const uint8_t *bits[4] = {k_dc_lum_bits, p->huff == 1 ? k_ac_long_bits : k_ac_lum_bits, k_dc_chr_bits,
                            p->huff == 1 ? k_ac_long_bits_chr : k_ac_chr_bits};
  const uint8_t *vals[4] = {k_dc_vals, k_ac_lum_vals, k_dc_vals, k_ac_chr_vals};
  const uint8_t cls[4] = {0x00, 0x10, 0x01, 0x11};
  for (int t = 0; t < (C == 1 ? 2 : 4); t++) {
    int n = 0;
    for (int i = 0; i < 16; i++) n += bits[t][i];
    raw(w, 0xFF); raw(w, 0xC4); raw16(w, 2 + 1 + 16 + n); raw(w, cls[t]);
    for (int i = 0; i < 16; i++) raw(w, bits[t][i]);
    for (int i = 0; i < n; i++) raw(w, vals[t][i]);
  }
  if (p->restart_interval > 0) {
    raw(w, 0xFF); raw(w, 0xDD); raw16(w, 4); raw16(w, p->restart_interval);
  }
  if (p->noninterleaved && C == 3) return;  /* the scans' headers come with the scans */
  /* SOS */
  raw(w, 0xFF); raw(w, 0xDA); raw16(w, 6 + 2 * C); raw(w, (uint8_t)C);
  for (int c = 0; c < C; c++) { raw(w, (uint8_t)(c + 1)); raw(w, c == 0 ? 0x00 : 0x11); }
  raw(w, 0); raw(w, 63); raw(w, 0);
}

/* One block: forward DCT, quantisation, DC clamp to an encodable difference,
 * Huffman coding with (dc, ac); returns the block's quantised coefficients in
 * qc and updates the predictor. */
static void code_block(bw_t *w, const uint8_t *plane, int pw, int x0, int y0, const uint16_t qnat[64],
                       const ehuff_t *dc, const ehuff_t *ac, int *pred, int16_t qc[64]) {
  float blk[64];
  for (int y = 0; y < 8; y++)
    for (int x = 0; x < 8; x++) blk[y * 8 + x] = (float)plane[(size_t)(y0 + y) * pw + x0 + x] - 128.0f;
  fdct_quant(blk, qnat, qc);
  int dcd = qc[0] - *pred;
  if (dcd > 2047 || dcd < -2047) { /* keep the DC difference encodable */
    qc[0] = (int16_t)(*pred + (dcd > 0 ? 2047 : -2047));
    dcd = qc[0] - *pred;
  }
  *pred = qc[0];
  int s = nbits_of(dcd);
  bw_put(w, dc->code[s], dc->len[s]);
  bw_put(w, (uint32_t)(dcd < 0 ? dcd - 1 : dcd), s);
  int run = 0;
  for (int k = 1; k < 64; k++) {
    int v = qc[k_zz[k]];
    if (!v) { run++; continue; }
    while (run > 15) { bw_put(w, ac->code[0xF0], ac->len[0xF0]); run -= 16; }
    int sz = nbits_of(v);
    int sym = (run << 4) | sz;
    bw_put(w, ac->code[sym], ac->len[sym]);
    bw_put(w, (uint32_t)(v < 0 ? v - 1 : v), sz);
    run = 0;
  }
  if (run) bw_put(w, ac->code[0], ac->len[0]);
}

/* RSTn between two restart intervals: byte-align with 1 bits, the marker. */
static void put_rst(bw_t *w, int *rst_idx) {
  bw_flush_ones(w);
  raw(w, 0xFF); raw(w, (uint8_t)(0xD0 | *rst_idx));
  w->raw_bits += 16; /* the unstuffed stream keeps RSTn as two bytes */
  *rst_idx = (*rst_idx + 1) & 7;
}

static inline uint8_t clampu8(double v) {
  long i = lround(v);
  return (uint8_t)(i < 0 ? 0 : i > 255 ? 255 : i);
}

long synth_encode(const synth_params *p, const uint8_t *pix, uint8_t *out, long cap,
                  int16_t *coefs, uint32_t *block_bits) {
  const int W = p->width, H = p->height, C = p->ncomp == 1 ? 1 : 3;
  int hs[3] = {1, 1, 1}, vs[3] = {1, 1, 1};
  if (C == 3) { hs[0] = p->h_samp; vs[0] = p->v_samp; }
  if (W <= 0 || H <= 0 || W > 65535 || H > 65535) return -1;
  /* luma factors 1, 2 or 4 (powers of two, as the reference accepts,
     src/format.cu:274-276), at most 10 blocks per MCU with the chroma */
  if ((hs[0] != 1 && hs[0] != 2 && hs[0] != 4) || (vs[0] != 1 && vs[0] != 2 && vs[0] != 4) || hs[0] * vs[0] > 8)
    return -1;
  cos_init();
  const int hmax = hs[0], vmax = vs[0];
  const int nbx = (W + 8 * hmax - 1) / (8 * hmax), nby = (H + 8 * vmax - 1) / (8 * vmax);
  uint16_t q[3][64];
  const int qmax = p->dqt16 ? 65535 : 255;
  synth_quant_table16(p->quality, 0, qmax, q[0]);
  synth_quant_table16(p->quality, 1, qmax, q[1]);
  synth_quant_table16(p->quality > 20 ? p->quality - 15 : p->quality + 15, 1, qmax, q[2]);

  /* component planes, padded to whole MCUs by edge replication */
  uint8_t *plane[3] = {0, 0, 0};
  int pw[3], ph[3];
  for (int c = 0; c < C; c++) {
    pw[c] = nbx * hs[c] * 8; ph[c] = nby * vs[c] * 8;
    plane[c] = (uint8_t *)malloc((size_t)pw[c] * ph[c]);
    if (!plane[c]) { for (int k = 0; k < c; k++) free(plane[k]); return -2; }
  }
  if (C == 1) {
    for (int y = 0; y < ph[0]; y++) {
      int sy = y < H ? y : H - 1;
      for (int x = 0; x < pw[0]; x++) plane[0][(size_t)y * pw[0] + x] = pix[(size_t)sy * W + (x < W ? x : W - 1)];
    }
  } else {
    /* full-resolution YCbCr (JFIF), then box-downsample chroma */
    uint8_t *ycc = (uint8_t *)malloc((size_t)W * H * 3);
    if (!ycc) { for (int k = 0; k < C; k++) free(plane[k]); return -2; }
    for (size_t i = 0; i < (size_t)W * H; i++) {
      double R = pix[3 * i], G = pix[3 * i + 1], B = pix[3 * i + 2];
      ycc[3 * i] = clampu8(0.299 * R + 0.587 * G + 0.114 * B);
      ycc[3 * i + 1] = clampu8(-0.168736 * R - 0.331264 * G + 0.5 * B + 128);
      ycc[3 * i + 2] = clampu8(0.5 * R - 0.418688 * G - 0.081312 * B + 128);
    }
    for (int c = 0; c < 3; c++) {
      const int fx = hmax / hs[c], fy = vmax / vs[c];
      const int cw = (W * hs[c] + hmax - 1) / hmax, chh = (H * vs[c] + vmax - 1) / vmax;
      for (int y = 0; y < ph[c]; y++) {
        int yy = y < chh ? y : chh - 1;
        for (int x = 0; x < pw[c]; x++) {
          int xx = x < cw ? x : cw - 1;
          int s = 0, n = 0;
          for (int dy = 0; dy < fy; dy++)
            for (int dx = 0; dx < fx; dx++) {
              int X = xx * fx + dx, Y = yy * fy + dy;
              if (X >= W) X = W - 1;
              if (Y >= H) Y = H - 1;
              s += ycc[((size_t)Y * W + X) * 3 + c];
              n++;
            }
          plane[c][(size_t)y * pw[c] + x] = (uint8_t)((s + n / 2) / n);
        }
      }
    }
    free(ycc);
  }

  ehuff_t dc_t[2], ac_t[2];
  build_ehuff(k_dc_lum_bits, k_dc_vals, &dc_t[0]);
  build_ehuff(p->huff == 1 ? k_ac_long_bits : k_ac_lum_bits, k_ac_lum_vals, &ac_t[0]);
  build_ehuff(k_dc_chr_bits, k_dc_vals, &dc_t[1]);
  build_ehuff(p->huff == 1 ? k_ac_long_bits_chr : k_ac_chr_bits, k_ac_chr_vals, &ac_t[1]);

  bw_t w = {out, cap, 0, 0, 0, 0, 0};
  put_header(&w, p, q);
  int pred[3] = {0, 0, 0};
  int rst_idx = 0;
  const long n_mcu = (long)nbx * nby;
  if (p->noninterleaved && C == 3) {
    /* one scan per component, blocks in raster order over the component's
       own grid: ceil(comp_width / 8) x ceil(comp_height / 8) blocks, fewer
       than the MCU layout where MCUs pad it (T.81 A.2.2) */
    const int bpm = hs[0] * vs[0] + 2;
    if (coefs) memset(coefs, 0, sizeof(int16_t) * 64 * (size_t)n_mcu * bpm);
    int slot0 = 0;
    for (int c = 0; c < 3; c++) {
      if (c == 1) {
        /* table 0 becomes the chroma tables for the chroma scans */
        const uint8_t *cb[2] = {k_dc_chr_bits, p->huff == 1 ? k_ac_long_bits_chr : k_ac_chr_bits};
        const uint8_t *cv[2] = {k_dc_vals, k_ac_chr_vals};
        for (int t = 0; t < 2; t++) {
          int n = 0;
          for (int i = 0; i < 16; i++) n += cb[t][i];
          raw(&w, 0xFF); raw(&w, 0xC4); raw16(&w, 2 + 1 + 16 + n); raw(&w, (uint8_t)(t << 4));
          for (int i = 0; i < 16; i++) raw(&w, cb[t][i]);
          for (int i = 0; i < n; i++) raw(&w, cv[t][i]);
        }
        if (p->restart_interval > 0) {
          raw(&w, 0xFF); raw(&w, 0xDD); raw16(&w, 4); raw16(&w, p->restart_interval + 1);
        }
      }
      /* mode 2: table 1 redefined before the Cr scan; Cb keeps the table it had */
      const int redefine = c == 2 && p->noninterleaved == 2 && p->qtables == 2;
      if (redefine) {
        if (p->dqt16) {
          raw(&w, 0xFF); raw(&w, 0xDB); raw16(&w, 131); raw(&w, 0x11);
          for (int k = 0; k < 64; k++) raw16(&w, q[2][k_zz[k]]);
        } else {
          raw(&w, 0xFF); raw(&w, 0xDB); raw16(&w, 67); raw(&w, 1);
          for (int k = 0; k < 64; k++) raw(&w, (uint8_t)q[2][k_zz[k]]);
        }
      }
      raw(&w, 0xFF); raw(&w, 0xDA); raw16(&w, 8); raw(&w, 1);
      raw(&w, (uint8_t)(c + 1)); raw(&w, 0x00);
      raw(&w, 0); raw(&w, 63); raw(&w, 0);
      const int ri = p->restart_interval > 0 ? p->restart_interval + (c > 0 ? 1 : 0) : 0;
      const int t = c == 0 ? 0 : 1, tq = redefine ? 2 : comp_qtab(p, c);
      const int cw = (W * hs[c] + hmax - 1) / hmax, chh = (H * vs[c] + vmax - 1) / vmax;
      const int bw = (cw + 7) / 8, bh = (chh + 7) / 8;
      int pr = 0, count = 0;
      rst_idx = 0;
      for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
          int16_t qc[64];
          code_block(&w, plane[c], pw[c], bx * 8, by * 8, q[tq], &dc_t[t], &ac_t[t], &pr, qc);
          const long mcu = (long)(by / vs[c]) * nbx + bx / hs[c];
          const long idx = mcu * bpm + slot0 + (by % vs[c]) * hs[c] + bx % hs[c];
          if (coefs) memcpy(coefs + idx * 64, qc, sizeof(qc));
          if (block_bits) block_bits[idx] = (uint32_t)w.raw_bits;
          if (ri > 0 && ++count == ri && (long)by * bw + bx + 1 < (long)bw * bh) {
            put_rst(&w, &rst_idx);
            count = 0;
            pr = 0;
          }
        }
      bw_flush_ones(&w);
      slot0 += hs[c] * vs[c];
    }
  } else {
    long blk_index = 0;
    int rst_count = 0;
    for (long m = 0; m < n_mcu; m++) {
      const int mx = (int)(m % nbx), my = (int)(m / nbx);
      for (int c = 0; c < C; c++) {
        const int t = c == 0 ? 0 : 1, tq = comp_qtab(p, c);
        for (int sy = 0; sy < vs[c]; sy++)
          for (int sx = 0; sx < hs[c]; sx++) {
            if (block_bits) block_bits[blk_index] = (uint32_t)w.raw_bits;
            int16_t qc[64];
            code_block(&w, plane[c], pw[c], (mx * hs[c] + sx) * 8, (my * vs[c] + sy) * 8, q[tq], &dc_t[t],
                       &ac_t[t], &pred[c], qc);
            if (coefs) memcpy(coefs + blk_index * 64, qc, sizeof(qc));
            blk_index++;
          }
      }
      if (p->restart_interval > 0 && ++rst_count == p->restart_interval && m + 1 < n_mcu) {
        put_rst(&w, &rst_idx);
        rst_count = 0;
        pred[0] = pred[1] = pred[2] = 0;
      }
    }
  }
  bw_flush_ones(&w);
  raw(&w, 0xFF); raw(&w, 0xD9);
  for (int c = 0; c < C; c++) free(plane[c]);
  if (w.overflow) return -3;
  return w.n;
}

long synth_make(const synth_params *p, uint8_t *out, long cap, int16_t *coefs, uint32_t *block_bits) {
  const int C = p->ncomp == 1 ? 1 : 3;
  uint8_t *pix = (uint8_t *)malloc((size_t)p->width * p->height * C);
  if (!pix) return -2;
  int e = synth_pixels(p, pix);
  long n = e ? e : synth_encode(p, pix, out, cap, coefs, block_bits);
  free(pix);
  return n;
}

long synth_num_blocks(const synth_params *p) {
  const int C = p->ncomp == 1 ? 1 : 3;
  const int hmax = C == 1 ? 1 : p->h_samp, vmax = C == 1 ? 1 : p->v_samp;
  const long nbx = (p->width + 8 * hmax - 1) / (8 * hmax), nby = (p->height + 8 * vmax - 1) / (8 * vmax);
  return nbx * nby * (hmax * vmax + (C == 3 ? 2 : 0));
}
