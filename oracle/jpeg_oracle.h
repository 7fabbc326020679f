/*
 * CPU oracle for the baseline-JPEG scan-decode path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference decoder's algorithm, used only as the
 * checker by tests, smoke() and the bench's CPU-baseline leg.  The product
 * decoder never calls it.  Each function cites the reference file:line it
 * follows.
 */
#ifndef JPEG_ORACLE_H
#define JPEG_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* error codes: include/format.h:7-14 */
#define ORC_NO_ERROR 0
#define ORC_SYNTAX_ERROR 1
#define ORC_UNSUPPORTED_ERROR 2
#define ORC_OOM_ERROR 3
#define ORC_FILE_ERROR 5

/* most scans of a file with several (progressive files: T.81 sets no limit;
 * libjpeg's scripts use 6 to 10) */
#define ORC_MAX_SCANS 64

typedef struct {
  int id, samples_x, samples_y, dq_id, dc_id, ac_id;
  int width, height, stride, plane_height;
} orc_channel;

typedef struct {
  int error, error_line;
  int width, height, num_channels;
  int num_blocks_x, num_blocks_y;   /* MCUs (reference naming) */
  int block_size_x, block_size_y;
  int restart_interval;
  int blocks_per_mcu;
  long num_blocks;                  /* 8x8 blocks in the scan */
  orc_channel ch[3];
  uint8_t dq[4][64];                /* as stored: zigzag order */
  uint8_t dht_bits[4][16];
  uint8_t dht_vals[4][256];
  int dht_present[4];
  /* entropy-coded segment (stuffed), from after the SOS header to EOI */
  const uint8_t *scan;
  size_t scan_len;
  int have_scan;
  int ext;                          /* extensions accepted (ORC_EXT_*), set by orc_parse_ext */
  uint16_t dq16[4][64];             /* the quant tables with 16-bit entries, zigzag order */
  /* ORC_EXT_NONINTERLEAVED: one scan per component (T.81 A.2.2); with
   * ORC_EXT_PROGRESSIVE the scans of a progressive frame; 0 for the single
   * interleaved scan above */
  int num_scans;
  struct {
    int comp, dc_id, ac_id, restart_interval;  /* the scan's (first) component and the state at its SOS */
    const uint8_t *scan;                       /* stuffed, from after its SOS header to the next marker */
    size_t scan_len;
    uint8_t bits[4][16], vals[4][256];         /* Huffman tables in effect at the scan */
    /* progressive scans: components (frame indexes, frame order) with their
     * DC / AC table ids, spectral band Ss..Se and successive approximation */
    int ncomp, comps[3], dcs[3], acs[3], ss, se, ah, al;
  } scans[ORC_MAX_SCANS];
  uint16_t comp_dq16[3][64];        /* each component's quant table when its first scan began (zigzag) */
  int progressive;                  /* SOF2 frame (ORC_EXT_PROGRESSIVE) */
} orc_jpeg;

/* Extensions beyond the reference: 16-bit quantisation tables (DQT Pq = 1,
 * T.81 B.2.4.1), which src/format.cu:432 rejects. */
#define ORC_EXT_DQT16 1
/* Non-interleaved scans (one component per scan, T.81 A.2.2-A.2.3), which
 * src/decodeScanGPU.cu:131 rejects: the scans may be separated by DHT, DQT
 * and DRI, and each uses the tables and restart interval in effect at its
 * SOS; each component's quant table is the one in effect when its scan
 * begins.  Coefficients come out in the interleaved (MCU) layout, with zero
 * blocks where MCUs pad a component's own block grid. */
#define ORC_EXT_NONINTERLEAVED 2
/* Progressive frames (SOF2, T.81 G.1), which src/format.cu:198-200 rejects:
 * DC first and refinement scans (interleaved or not), AC first scans with
 * end-of-band runs and AC refinement scans with correction bits, each with the
 * tables and restart interval in effect at its SOS.  The progression must
 * follow T.81 G.1.1.1: a band's first scan has Ah = 0, each later scan of it
 * Ah = the previous Al and Al = Ah - 1, a component's AC scans come after its
 * DC first scan; anything else is SYNTAX_ERROR.  Coefficients come out in the
 * interleaved (MCU) layout; each component is dequantised with the table in
 * effect at its first scan. */
#define ORC_EXT_PROGRESSIVE 4

/* Parse markers, SOF, DHT, DQT, DRI and the SOS header of an in-memory file
 * (src/format.cu:113-215, :242-442; src/decodeScanGPU.cu:124-141).
 * Returns the reference error code. */
int orc_parse(const uint8_t *file, size_t size, orc_jpeg *j);

/* Remove byte stuffing, keeping RSTn (src/decodeScanGPU.cu:11-24).
 * out must hold scan_len bytes; returns the unstuffed length. */
size_t orc_unstuff(const uint8_t *scan, size_t len, uint8_t *out);

/* Serial scan decode (src/decodeScanGPU.cu:28-100,435-490 ==
 * src/decodeScanCPU.cu:10-152).  All outputs optional:
 *  qcoef  int16 [num_blocks*64]  quantised coefficients, natural order, MCU order
 *  work   int32 per component    dequantised working_space in the reference layout
 *         (work[c] has stride*plane_height entries, 64 per block in decode order)
 *  bitpos uint32 [num_blocks]    start bit of each block in the unstuffed segment
 *                                (interleaved scans)
 * Non-interleaved scans (num_scans > 0) are decoded one after the other, each
 * block written where the interleaved scan holds it; progressive scans
 * likewise, into coefficients that start at zero (bitpos is left untouched).
 * Returns the reference error code. */
int orc_decode_scan(const orc_jpeg *j, int16_t *qcoef, int32_t *work[3], uint32_t *bitpos);

/* Integer Chen-Wang iDCT of one dequantised block into an 8x8 pixel tile
 * (CPUonly/src/pixelTransformCPU.c:14-107). */
void orc_idct_int(const int32_t in[64], uint8_t *out, int stride);
/* Same integer algorithm without the all-zero shortcuts
 * (src/pixelTransformGPU.cu:25-142). */
void orc_idct_int_noshortcut(const int32_t in[64], uint8_t *out, int stride);
/* Exact T.81 A.3.3 inverse DCT in double precision, rounded and clamped. */
void orc_idct_f64(const int32_t in[64], uint8_t *out, int stride);
/* Unrounded double-precision samples (level shift included). */
void orc_idct_f64_raw(const int32_t in[64], double out[64]);

/* Planar component planes from quantised coefficients (scan order).
 * planes[c] holds stride*plane_height bytes (src/format.cu:295-303,
 * src/pixelTransformGPU.cu:120-131).  mode 0 = integer iDCT, 1 = FP64 iDCT. */
int orc_planes(const orc_jpeg *j, const int16_t *qcoef, uint8_t *planes[3], int mode);

/* Nearest chroma upsample + integer YCbCr->RGB, or gray stride removal
 * (CPUonly/src/pixelTransformCPU.c:110-172).  out: W*H*3 or W*H. */
int orc_colour(const orc_jpeg *j, uint8_t *const planes[3], uint8_t *out);

/* Whole decode of an in-memory file (parse, scan, iDCT, colour). */
int orc_decode_file(const uint8_t *file, size_t size, int idct_mode, uint8_t *out,
                    size_t out_cap, int *width, int *height, int *ncomp);
/* orc_parse / orc_decode_file with the extensions `ext` accepted. */
int orc_parse_ext(const uint8_t *file, size_t size, int ext, orc_jpeg *j);
int orc_decode_file_ext(const uint8_t *file, size_t size, int ext, int idct_mode, uint8_t *out,
                        size_t out_cap, int *width, int *height, int *ncomp);

#ifdef __cplusplus
}
#endif
#endif
