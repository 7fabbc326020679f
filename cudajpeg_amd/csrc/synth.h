/* Seeded synthetic-image generator and baseline JPEG encoder (C ABI). */
#ifndef CUDAJPEG_AMD_SYNTH_H
#define CUDAJPEG_AMD_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint64_t seed;
  int32_t width, height;
  int32_t ncomp;            /* 1 (grayscale) or 3 (YCbCr) */
  int32_t h_samp, v_samp;   /* luma sampling factors; chroma is 1x1 */
  int32_t quality;          /* IJG quality 1..100 */
  int32_t restart_interval; /* MCUs per restart interval, 0 = none */
  double sigma;             /* Gaussian noise std */
  double amp_total;         /* sum of the 6 sinusoid amplitudes */
  double random_tile_frac;  /* fraction of 64x64 tiles replaced by uniform noise */
  int32_t qtables;          /* quant tables for YCbCr: 2 (Y | CbCr) or 3 (Y | Cb | Cr) */
  int32_t huff;             /* AC Huffman tables: 0 Annex K, 1 long codes (60 of 11 bits, 92 of 16) */
  int32_t dqt16;            /* 1: quant tables written with 16-bit entries (Pq = 1), values up to 65535 */
  int32_t noninterleaved;   /* 1 (YCbCr): one scan per component, each in its component's block raster
                               order (T.81 A.2.2); a DHT between the luma and chroma scans redefines
                               table 0 with the chroma tables, which the chroma scans use, and with
                               restarts a DRI there sets the chroma scans' interval to the luma's + 1;
                               2: also, with qtables = 2, a DQT before the Cr scan redefines table 1
                               as the third table, which Cr is quantised with (the same quantisation
                               as qtables = 3) */
} synth_params;

/* W*H*ncomp interleaved pixels; returns 0 on success. */
int synth_pixels(const synth_params *p, uint8_t *out);
/* Encodes pixels; returns the file size, or <0 on error (-3 = cap too small).
 * coefs (optional): int16[num_blocks*64], quantised, natural order, in MCU
 * order (the interleaved scan's order; with noninterleaved, the blocks no scan
 * codes -- the MCU padding outside a component's block grid -- are zero).
 * block_bits (optional, interleaved scans): uint32[num_blocks], start bit of
 * every block in the unstuffed entropy-coded segment. */
long synth_encode(const synth_params *p, const uint8_t *pixels, uint8_t *out, long cap,
                  int16_t *coefs, uint32_t *block_bits);
long synth_make(const synth_params *p, uint8_t *out, long cap, int16_t *coefs, uint32_t *block_bits);
long synth_num_blocks(const synth_params *p);
void synth_quant_table(int quality, int chroma, uint8_t out_natural[64]);
/* The same scaled table clamped to [1, maxv] (255 for 8-bit tables). */
void synth_quant_table16(int quality, int chroma, int maxv, uint16_t out_natural[64]);

#ifdef __cplusplus
}
#endif
#endif
