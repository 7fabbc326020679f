/*
 * CPU oracle for the baseline-JPEG scan-decode path -- TEST INFRASTRUCTURE.
 *
 * Restates the reference algorithm (jndean/CUDAjpeg) in plain C so that the
 * HIP decoder can be checked against it.  Only tests, smoke() and the bench's
 * CPU-baseline leg use this file; the product decoder never does.
 *
 * Pinned by: (1) the encoder's own emitted coefficients and block offsets,
 * (2) the reference sources compiled from /root/reference into oracle/_ref
 * (CPUonly pixels; src/decodeScanCPU.cu dequantised coefficients) on the
 * golden files, (3) closed forms (iDCT shortcut algebra, FP64 iDCT vs the
 * integer iDCT bound), (4) for progressive scans (T.81 G.1.2, which the
 * reference does not decode): libjpeg's baseline and progressive files of the
 * same pixels, whose coefficients must agree, and every scan prefix of a
 * progressive file against T.81's point transform of the final coefficients.
 */
#include "jpeg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* include/format.h:117-120 */
static const uint8_t kDeZigZag[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15,
    23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* CPUonly/include/pixelTransformCPU.h:5-10 */
#define W1 2841
#define W2 2676
#define W3 2408
#define W5 1609
#define W6 1108
#define W7 565

#define FAIL(e) do { j->error = (e); j->error_line = __LINE__; return (e); } while (0)

static unsigned read16(const uint8_t *p) { return (unsigned)(p[0] << 8 | p[1]); }

/* ---------------------------------------------------------------- parse -- */
/* src/format.cu:242-354 */
static int parse_sof(orc_jpeg *j, const uint8_t *block, const uint8_t *end) {
  unsigned block_len = read16(block);
  if (block_len < 9 || block + block_len >= end) FAIL(ORC_SYNTAX_ERROR);
  if (block[2] != 8) FAIL(ORC_UNSUPPORTED_ERROR);
  j->height = (int)read16(block + 3);
  j->width = (int)read16(block + 5);
  if (!j->width || !j->height) FAIL(ORC_SYNTAX_ERROR);
  j->num_channels = block[7];
  if (j->num_channels != 1 && j->num_channels != 3) FAIL(ORC_UNSUPPORTED_ERROR);
  if (block_len < 8u + (unsigned)(j->num_channels * 3)) FAIL(ORC_SYNTAX_ERROR);
  block += 8;
  int sxm = 0, sym = 0;
  for (int i = 0; i < j->num_channels; i++, block += 3) {
    orc_channel *c = &j->ch[i];
    c->id = block[0];
    c->samples_x = block[1] >> 4;
    c->samples_y = block[1] & 0xF;
    c->dq_id = block[2];
    if (!c->samples_x || !c->samples_y || c->dq_id > 3) FAIL(ORC_SYNTAX_ERROR);
    if ((c->samples_x & (c->samples_x - 1)) || (c->samples_y & (c->samples_y - 1)))
      FAIL(ORC_UNSUPPORTED_ERROR);
    if (c->samples_x > sxm) sxm = c->samples_x;
    if (c->samples_y > sym) sym = c->samples_y;
  }
  if (j->num_channels == 1) { /* src/format.cu:281-284 */
    j->ch[0].samples_x = sxm = 1;
    j->ch[0].samples_y = sym = 1;
  }
  j->block_size_x = sxm << 3;
  j->block_size_y = sym << 3;
  j->num_blocks_x = (j->width + j->block_size_x - 1) / j->block_size_x;
  j->num_blocks_y = (j->height + j->block_size_y - 1) / j->block_size_y;
  j->blocks_per_mcu = 0;
  for (int i = 0; i < j->num_channels; i++) {
    orc_channel *c = &j->ch[i];
    c->width = (j->width * c->samples_x + sxm - 1) / sxm;
    c->height = (j->height * c->samples_y + sym - 1) / sym;
    c->stride = j->num_blocks_x * (c->samples_x << 3);
    c->plane_height = j->num_blocks_y * (c->samples_y << 3);
    if ((c->width < 3 && c->samples_x != sxm) || (c->height < 3 && c->samples_y != sym))
      FAIL(ORC_UNSUPPORTED_ERROR);
    j->blocks_per_mcu += c->samples_x * c->samples_y;
  }
  if (j->num_channels == 3) { /* src/format.cu:346-350 */
    if (j->ch[1].width != j->ch[2].width || j->ch[1].height != j->ch[2].height ||
        j->ch[0].width != j->width || j->ch[0].height != j->height)
      FAIL(ORC_UNSUPPORTED_ERROR);
  }
  j->num_blocks = (long)j->num_blocks_x * j->num_blocks_y * j->blocks_per_mcu;
  return 0;
}

/* src/format.cu:357-400 (table_id = Tc*2 + Th; the 65536-entry LUT itself is
 * built at decode time from the stored BITS/HUFFVAL, see build_lut) */
static int parse_dht(orc_jpeg *j, const uint8_t *pos, const uint8_t *end) {
  unsigned block_len = read16(pos);
  const uint8_t *block_end = pos + block_len;
  if (block_end > end) FAIL(ORC_SYNTAX_ERROR);
  pos += 2;
  while (pos < block_end) {
    uint8_t val = pos[0];
    if (val & 0xEC) FAIL(ORC_SYNTAX_ERROR);
    if (val & 0x02) FAIL(ORC_UNSUPPORTED_ERROR);
    int id = (val | (val >> 3)) & 3;
    const uint8_t *tuple = pos + 17;
    int remain = 65536, n = 0;
    for (int l = 1; l <= 16; l++) {
      int count = pos[l];
      j->dht_bits[id][l - 1] = (uint8_t)count;
      if (!count) continue;
      if (tuple + count > block_end) FAIL(ORC_SYNTAX_ERROR);
      remain -= count << (16 - l);
      if (remain < 0) FAIL(ORC_SYNTAX_ERROR);
      if (n + count > 256) FAIL(ORC_SYNTAX_ERROR); /* HUFFVAL holds at most 256 symbols */
      for (int i = 0; i < count; i++) j->dht_vals[id][n++] = *tuple++;
    }
    j->dht_present[id] = 1;
    pos = tuple;
  }
  if (pos != block_end) FAIL(ORC_SYNTAX_ERROR);
  return 0;
}

/* src/format.cu:424-442; with ORC_EXT_DQT16 also tables of 16-bit entries
 * (T.81 B.2.4.1: Pq = 1, 64 big-endian values).  dq keeps the reference's
 * 8-bit view (entries clamped to 255), dq16 the values. */
static int parse_dqt(orc_jpeg *j, const uint8_t *p, const uint8_t *end, const uint8_t **out_end) {
  unsigned block_len = read16(p);
  const uint8_t *block_end = p + block_len;
  if (block_end > end) FAIL(ORC_SYNTAX_ERROR);
  const uint8_t *pos = p + 2;
  while (pos + 65 <= block_end) {
    uint8_t id = pos[0];
    int pq = id >> 4;
    if (pq == 1 && (j->ext & ORC_EXT_DQT16)) {
      if ((id & 0x0C) || pos + 129 > block_end) FAIL(ORC_SYNTAX_ERROR);
      for (int k = 0; k < 64; k++) {
        unsigned v = (unsigned)pos[1 + 2 * k] << 8 | pos[2 + 2 * k];
        j->dq16[id & 3][k] = (uint16_t)v;
        j->dq[id & 3][k] = (uint8_t)(v > 255 ? 255 : v);
      }
      pos += 129;
      continue;
    }
    if (id & 0xFC) FAIL(ORC_SYNTAX_ERROR);
    memcpy(j->dq[id], pos + 1, 64);
    for (int k = 0; k < 64; k++) j->dq16[id][k] = pos[1 + k];
    pos += 65;
  }
  if (pos != block_end) FAIL(ORC_SYNTAX_ERROR);
  *out_end = block_end;
  return 0;
}

/* src/decodeScanGPU.cu:124-141 */
static int parse_sos(orc_jpeg *j, const uint8_t *p, const uint8_t *end, const uint8_t **after) {
  unsigned header_len = read16(p);
  if (p + header_len > end) FAIL(ORC_SYNTAX_ERROR);
  const uint8_t *pos = p + 2;
  if (header_len < 4u + 2u * (unsigned)j->num_channels) FAIL(ORC_SYNTAX_ERROR);
  if (*(pos++) != j->num_channels) FAIL(ORC_UNSUPPORTED_ERROR);
  for (int i = 0; i < j->num_channels; i++, pos += 2) {
    if (pos[0] != j->ch[i].id) FAIL(ORC_SYNTAX_ERROR);
    if (pos[1] & 0xEE) FAIL(ORC_SYNTAX_ERROR);
    j->ch[i].dc_id = pos[1] >> 4;
    j->ch[i].ac_id = (pos[1] & 1) | 2;
  }
  if (pos[0] || pos[1] != 63 || pos[2]) FAIL(ORC_UNSUPPORTED_ERROR);
  *after = p + header_len;
  return 0;
}

static const uint8_t *scan_end(const uint8_t *d, const uint8_t *end);

/* ORC_EXT_NONINTERLEAVED: SOS of a one-component scan (T.81 B.2.3 with
 * Ns = 1), which src/decodeScanGPU.cu:124-131 rejects.  Records the scan
 * with the tables and restart interval in effect, latches the component's
 * quant table, and finds the scan's end: the first 0xFF followed by neither
 * 0x00 nor RSTn (T.81 B.1.1.2 and B.1.1.5), at the latest the EOI. */
static int parse_sos_ni(orc_jpeg *j, const uint8_t *p, const uint8_t *end, const uint8_t **next) {
  unsigned header_len = read16(p);
  if (header_len != 8 || p + header_len > end) FAIL(ORC_SYNTAX_ERROR);
  const uint8_t *pos = p + 3;
  int c = 0;
  while (c < j->num_channels && j->ch[c].id != pos[0]) c++;
  if (c == j->num_channels) FAIL(ORC_SYNTAX_ERROR);
  for (int k = 0; k < j->num_scans; k++)
    if (j->scans[k].comp == c) FAIL(ORC_SYNTAX_ERROR); /* each component in one scan */
  if (j->num_scans >= j->num_channels) FAIL(ORC_SYNTAX_ERROR);
  if (pos[1] & 0xEE) FAIL(ORC_SYNTAX_ERROR);
  if (pos[2] || pos[3] != 63 || pos[4]) FAIL(ORC_UNSUPPORTED_ERROR);
  const int k = j->num_scans++;
  j->scans[k].comp = c;
  j->scans[k].dc_id = pos[1] >> 4;
  j->scans[k].ac_id = (pos[1] & 1) | 2;
  j->scans[k].restart_interval = j->restart_interval;
  memcpy(j->scans[k].bits, j->dht_bits, sizeof(j->dht_bits));
  memcpy(j->scans[k].vals, j->dht_vals, sizeof(j->dht_vals));
  memcpy(j->comp_dq16[c], j->dq16[j->ch[c].dq_id], sizeof(j->comp_dq16[c]));
  const uint8_t *d = p + header_len, *e = scan_end(d, end);
  j->scans[k].scan = d;
  j->scans[k].scan_len = (size_t)(e - d);
  j->have_scan = 1;
  *next = e;
  return 0;
}

/* End of a scan's entropy-coded data: the first 0xFF followed by neither
 * 0x00 nor RSTn (T.81 B.1.1.2 and B.1.1.5), at the latest the EOI. */
static const uint8_t *scan_end(const uint8_t *d, const uint8_t *end) {
  const uint8_t *e = d;
  while (e < end && !(e[0] == 0xFF && e[1] != 0x00 && (e[1] & 0xF8) != 0xD0)) e++;
  return e;
}

/* ORC_EXT_PROGRESSIVE: SOS of a progressive frame (T.81 B.2.3, G.1.1.1).
 * Ns components in frame order; a DC scan (Ss = 0) has Se = 0, an AC scan
 * 1 <= Ss <= Se <= 63 and one component; Al <= 13 and a refinement scan
 * (Ah > 0) has Al = Ah - 1.  cbits[c][k] is the Al of the last scan that
 * coded coefficient k of component c (-1: none yet): a first scan needs -1,
 * a refinement Ah; and an AC scan needs the component's DC begun.  Records
 * the scan with the tables and restart interval in effect and latches the
 * quant table of each component at its first scan. */
static int parse_sos_prog(orc_jpeg *j, const uint8_t *p, const uint8_t *end, const uint8_t **next,
                          signed char cbits[3][64]) {
  if (!j->num_channels) FAIL(ORC_SYNTAX_ERROR);
  unsigned header_len = read16(p);
  if (p + header_len > end || header_len < 3) FAIL(ORC_SYNTAX_ERROR);
  const int ns = p[2];
  if (ns < 1 || ns > j->num_channels || header_len != 6u + 2u * (unsigned)ns) FAIL(ORC_SYNTAX_ERROR);
  if (j->num_scans >= ORC_MAX_SCANS) FAIL(ORC_UNSUPPORTED_ERROR);
  const int k = j->num_scans;
  const uint8_t *pos = p + 3;
  int last = -1;
  for (int i = 0; i < ns; i++, pos += 2) {
    int c = last + 1;
    while (c < j->num_channels && j->ch[c].id != pos[0]) c++;
    if (c == j->num_channels) FAIL(ORC_SYNTAX_ERROR); /* unknown, repeated or out of frame order */
    if (pos[1] & 0xEE) FAIL(ORC_SYNTAX_ERROR);
    j->scans[k].comps[i] = c;
    j->scans[k].dcs[i] = pos[1] >> 4;
    j->scans[k].acs[i] = (pos[1] & 1) | 2;
    last = c;
  }
  const int ss = pos[0], se = pos[1], ah = pos[2] >> 4, al = pos[2] & 15;
  if (ss == 0 ? se != 0 : (se < ss || se > 63 || ns != 1)) FAIL(ORC_SYNTAX_ERROR);
  if (al > 13 || (ah && al != ah - 1)) FAIL(ORC_SYNTAX_ERROR);
  for (int i = 0; i < ns; i++) {
    signed char *cb = cbits[j->scans[k].comps[i]];
    if (ss > 0 && cb[0] < 0) FAIL(ORC_SYNTAX_ERROR);
    for (int q = ss; q <= se; q++) {
      if (cb[q] != (ah ? ah : -1)) FAIL(ORC_SYNTAX_ERROR);
      cb[q] = (signed char)al;
    }
  }
  for (int i = 0; i < ns; i++) {
    const int c = j->scans[k].comps[i];
    int seen = 0;
    for (int t = 0; t < k; t++)
      for (int q = 0; q < j->scans[t].ncomp; q++) seen |= j->scans[t].comps[q] == c;
    if (!seen) memcpy(j->comp_dq16[c], j->dq16[j->ch[c].dq_id], sizeof(j->comp_dq16[c]));
  }
  j->num_scans++;
  j->scans[k].ncomp = ns;
  j->scans[k].comp = j->scans[k].comps[0];
  j->scans[k].dc_id = j->scans[k].dcs[0];
  j->scans[k].ac_id = j->scans[k].acs[0];
  j->scans[k].ss = ss;
  j->scans[k].se = se;
  j->scans[k].ah = ah;
  j->scans[k].al = al;
  j->scans[k].restart_interval = j->restart_interval;
  memcpy(j->scans[k].bits, j->dht_bits, sizeof(j->dht_bits));
  memcpy(j->scans[k].vals, j->dht_vals, sizeof(j->dht_vals));
  const uint8_t *d = p + header_len, *e = scan_end(d, end);
  j->scans[k].scan = d;
  j->scans[k].scan_len = (size_t)(e - d);
  j->have_scan = 1;
  *next = e;
  return 0;
}

/* src/format.cu:113-215 (file read replaced by an in-memory buffer) */
int orc_parse(const uint8_t *buf, size_t size, orc_jpeg *j) { return orc_parse_ext(buf, size, 0, j); }

int orc_parse_ext(const uint8_t *buf, size_t size, int ext, orc_jpeg *j) {
  memset(j, 0, sizeof(*j));
  j->ext = ext;
  if (size < 4 || buf[0] != 0xFF || buf[1] != 0xD8 || buf[size - 2] != 0xFF || buf[size - 1] != 0xD9)
    FAIL(ORC_FILE_ERROR);
  if (size < 6) FAIL(ORC_SYNTAX_ERROR);
  const uint8_t *end = buf + size - 2;
  const uint8_t *pos = buf + 2;
  signed char cbits[3][64];
  memset(cbits, -1, sizeof(cbits));
  for (;;) {
    if (pos > end) FAIL(ORC_SYNTAX_ERROR);
    if (pos[0] != 0xFF) FAIL(ORC_SYNTAX_ERROR);
    pos += 2;
    uint8_t m = pos[-1];
    int e = 0;
    switch (m) {
      case 0xC0:
        e = parse_sof(j, pos, end);
        if (!e) pos += read16(pos);
        break;
      case 0xC2: /* src/format.cu:198-200 rejects it as an unknown marker */
        if (!(j->ext & ORC_EXT_PROGRESSIVE)) FAIL(ORC_SYNTAX_ERROR);
        e = parse_sof(j, pos, end);
        if (!e) {
          j->progressive = 1;
          pos += read16(pos);
        }
        break;
      case 0xC4:
        e = parse_dht(j, pos, end);
        if (!e) pos += read16(pos);
        break;
      case 0xDB: {
        const uint8_t *ne;
        e = parse_dqt(j, pos, end, &ne);
        if (!e) pos = ne;
        break;
      }
      case 0xDD: { /* src/format.cu:404-421 */
        unsigned bl = read16(pos);
        if (bl < 2 || pos + bl > end) FAIL(ORC_SYNTAX_ERROR);
        j->restart_interval = (int)read16(pos + 2);
        pos += bl;
        break;
      }
      case 0xDA: {
        const uint8_t *after;
        if (j->progressive) {
          e = parse_sos_prog(j, pos, end, &after, cbits);
          if (!e) pos = after;
          break;
        }
        if ((j->ext & ORC_EXT_NONINTERLEAVED) && j->num_channels > 1 && read16(pos) >= 3 && pos[2] == 1) {
          e = parse_sos_ni(j, pos, end, &after);
          if (!e) pos = after;
          break;
        }
        if (j->num_scans) FAIL(ORC_SYNTAX_ERROR); /* an interleaved scan after one-component scans */
        e = parse_sos(j, pos, end, &after);
        if (!e) {
          /* the scan runs to the EOI (src/decodeScanGPU.cu:13 unstuffs up to jpg->end) */
          j->scan = after;
          j->scan_len = (size_t)(end - after);
          j->have_scan = 1;
          pos = end;
        }
        break;
      }
      case 0xFE:
        pos += read16(pos);
        break;
      case 0xD9:
        break;
      default:
        if ((m & 0xF0) == 0xE0) pos += read16(pos);
        else FAIL(ORC_SYNTAX_ERROR);
    }
    if (e) return e;
    if (m == 0xD9) break;
  }
  if (!j->progressive && j->num_scans && j->num_scans != j->num_channels)
    FAIL(ORC_SYNTAX_ERROR); /* a component never coded */
  if (j->progressive) { /* a component without scans keeps zero coefficients: any table will do */
    for (int c = 0; c < j->num_channels; c++)
      if (cbits[c][0] < 0) memcpy(j->comp_dq16[c], j->dq16[j->ch[c].dq_id], sizeof(j->comp_dq16[c]));
  }
  return 0;
}

/* src/decodeScanGPU.cu:11-24 */
size_t orc_unstuff(const uint8_t *src, size_t len, uint8_t *dst) {
  size_t s = 0, d = 0;
  while (s < len) {
    uint8_t b = src[s++];
    dst[d++] = b;
    if (b != 0xFF) continue;
    uint8_t nx = s < len ? src[s] : 0xFF; /* past the scan: the EOI's 0xFF */
    s++;
    if (nx == 0x00) continue;
    if ((nx & 0xF8) == 0xD0) dst[d++] = nx;
  }
  return d;
}

/* ----------------------------------------------------------- scan decode -- */
typedef struct { uint8_t tuple, num_bits; } vlc_t; /* include/format.h:30-33 */

/* The reference's 65536-entry direct LUT (src/format.cu:371-391): each code of
 * length L is spread over 2^(16-L) consecutive entries; the rest are invalid. */
static void build_lut(const uint8_t bits[16], const uint8_t *vals, vlc_t *vlc) {
  int remain = 65536, spread = 65536, k = 0;
  vlc_t *v = vlc;
  for (int l = 1; l <= 16; l++) {
    spread >>= 1;
    int count = bits[l - 1];
    remain -= count << (16 - l);
    for (int i = 0; i < count; i++, k++)
      for (int s = spread; s; s--, v++) { v->num_bits = (uint8_t)l; v->tuple = vals[k]; }
  }
  while (remain-- > 0) { v->num_bits = 0; v->tuple = 0; v++; }
}

typedef struct {
  const uint8_t *buf, *pos, *end;
  uint32_t bufbits;
  int num_bufbits;
  int error;
} reader_t;

/* src/decodeScanGPU.cu:28-37: 0xFF beyond the end */
static int show_bits(reader_t *r, int n) {
  if (!n) return 0;
  while (r->num_bufbits < n) {
    uint8_t b = (r->pos >= r->end) ? 0xFF : *r->pos++;
    r->bufbits = (r->bufbits << 8) | b;
    r->num_bufbits += 8;
  }
  return (int)((r->bufbits >> (r->num_bufbits - n)) & ((1u << n) - 1));
}
static int get_bits(reader_t *r, int n) {
  int v = show_bits(r, n);
  r->num_bufbits -= n;
  return v;
}
/* src/decodeScanGPU.cu:50-75 (EXTEND, T.81 F.2.2.1) */
static int get_vlc(reader_t *r, const vlc_t *tab, uint8_t *code) {
  int sym = show_bits(r, 16);
  vlc_t v = tab[sym];
  if (!v.num_bits) { r->error = ORC_SYNTAX_ERROR; return 0; }
  r->num_bufbits -= v.num_bits;
  if (code) *code = v.tuple;
  int nb = v.tuple & 0x0F;
  if (!nb) return 0;
  int value = get_bits(r, nb);
  if (value < (1 << (nb - 1))) value += ((-1) << nb) + 1;
  return value;
}

/* One block (src/decodeScanGPU.cu:78-100): DC difference and AC run/levels
 * into q (natural order) and, when w, the dequantised values. */
static int decode_block(reader_t *r, const vlc_t *dct, const vlc_t *act, const uint16_t *dq, int *pred,
                        int16_t q[64], int32_t *w) {
  memset(q, 0, 64 * sizeof(int16_t));
  *pred += get_vlc(r, dct, NULL);
  q[0] = (int16_t)*pred;
  if (w) { memset(w, 0, 64 * sizeof(int32_t)); w[0] = *pred * dq[0]; }
  int coef = 0;
  uint8_t code = 0;
  do {
    int value = get_vlc(r, act, &code);
    if (!code) break;
    if (!(code & 0x0F) && code != 0xF0) return ORC_SYNTAX_ERROR;
    coef += (code >> 4) + 1;
    if (coef > 63) return ORC_SYNTAX_ERROR;
    q[kDeZigZag[coef]] = (int16_t)value;
    if (w) w[kDeZigZag[coef]] = value * dq[coef];
  } while (coef < 63);
  return r->error;
}

/* ORC_EXT_NONINTERLEAVED: the one-component scans in file order.  Scan k
 * codes component c's own grid of ceil(width_c / 8) x ceil(height_c / 8)
 * blocks in raster order (T.81 A.2.2), a restart interval counting blocks;
 * block (bx, by) is block (by % v) h + bx % h of component c in MCU
 * (bx / h, by / v), and goes where the interleaved scan decodes it. */
static int decode_scans_ni(orc_jpeg *j, int16_t *qcoef, int32_t *work[3]) {
  const long bpm = j->blocks_per_mcu;
  if (qcoef) memset(qcoef, 0, sizeof(int16_t) * 64 * (size_t)j->num_blocks);
  for (int c = 0; c < j->num_channels; c++)
    if (work && work[c]) memset(work[c], 0, sizeof(int32_t) * (size_t)j->ch[c].stride * j->ch[c].plane_height);
  vlc_t *lut = (vlc_t *)malloc(sizeof(vlc_t) * 65536 * 2);
  if (!lut) FAIL(ORC_OOM_ERROR);
  int err = 0;
  for (int k = 0; k < j->num_scans && !err; k++) {
    const int c = j->scans[k].comp;
    const orc_channel *ch = &j->ch[c];
    int slot0 = 0;
    for (int i = 0; i < c; i++) slot0 += j->ch[i].samples_x * j->ch[i].samples_y;
    build_lut(j->scans[k].bits[j->scans[k].dc_id], j->scans[k].vals[j->scans[k].dc_id], lut);
    build_lut(j->scans[k].bits[j->scans[k].ac_id], j->scans[k].vals[j->scans[k].ac_id], lut + 65536);
    uint8_t *ub = (uint8_t *)malloc(j->scans[k].scan_len + 2);
    if (!ub) { err = ORC_OOM_ERROR; break; }
    size_t ulen = orc_unstuff(j->scans[k].scan, j->scans[k].scan_len, ub);
    reader_t r = {ub, ub, ub + ulen, 0, 0, 0};
    const int h = ch->samples_x, v = ch->samples_y;
    const int bw = (ch->width + 7) / 8, bh = (ch->height + 7) / 8;
    const int ri = j->scans[k].restart_interval;
    int pred = 0, rcount = ri, next_rst = 0;
    for (int by = 0; by < bh && !err; by++)
      for (int bx = 0; bx < bw && !err; bx++) {
        const long mcu = (long)(by / v) * j->num_blocks_x + bx / h;
        const long local = (long)(by % v) * h + bx % h;
        int16_t q[64];
        int32_t *w = work && work[c] ? work[c] + (mcu * h * v + local) * 64 : NULL;
        err = decode_block(&r, lut, lut + 65536, j->comp_dq16[c], &pred, q, w);
        if (!err && qcoef) memcpy(qcoef + (mcu * bpm + slot0 + local) * 64, q, sizeof(q));
        /* src/decodeScanGPU.cu:476-486, per block */
        if (!err && ri && !(--rcount) && r.pos < r.end) {
          r.num_bufbits &= 0xF8;
          int m = get_bits(&r, 16);
          if ((m & 0xFFF8) != 0xFFD0 || (m & 7) != next_rst) err = ORC_SYNTAX_ERROR;
          next_rst = (next_rst + 1) & 7;
          rcount = ri;
          pred = 0;
        }
      }
    free(ub);
  }
  free(lut);
  if (err) FAIL(err);
  return 0;
}

/* ORC_EXT_PROGRESSIVE (T.81 G.1.2).  Every scan adds to one coefficient
 * array in the interleaved layout, zero at the start.  A scan of several
 * components runs over the frame's MCUs (each component's h x v blocks, as in
 * the interleaved baseline scan), a one-component scan over the component's
 * own grid of ceil(width_c / 8) x ceil(height_c / 8) blocks (T.81 A.2.2);
 * the restart interval counts those units.  The four procedures:
 *  DC first   (Ss = 0, Ah = 0): DC difference as in baseline, stored << Al;
 *  DC refine  (Ss = 0, Ah > 0): one raw bit per block, ORed in at bit Al;
 *  AC first   (Ss > 0, Ah = 0): run/size symbols over Ss..Se, values << Al;
 *             EOBn (size 0, run r < 15) ends this block and the next
 *             2^r + (r extra bits) - 1 blocks of the band (G.1.2.2);
 *  AC refine  (Ss > 0, Ah > 0): symbols (run, size 0 or 1); a size-1 symbol
 *             brings a sign bit for a new coefficient +-2^Al placed at the
 *             (run + 1)-th coefficient of the band still zero; every nonzero
 *             coefficient passed on the way takes a correction bit, and a 1
 *             adds 2^Al to its magnitude; ZRL passes 16 zero coefficients;
 *             EOBn ends the symbols of this block and the next blocks of the
 *             run, whose nonzero coefficients still take correction bits
 *             (G.1.2.3).
 * A restart resets the DC predictors and the EOB run.  An invalid code, a
 * refinement size other than 0 or 1, or a run past Se is SYNTAX_ERROR. */
typedef struct {
  reader_t r;
  const vlc_t *dct[3], *act;
  int pred[3], eobrun;
  int ss, se, ah, al;
} prog_t;

static int prog_block(prog_t *P, int ci, int16_t *q) {
  reader_t *r = &P->r;
  uint8_t code = 0;
  if (P->ss == 0) {
    if (P->ah == 0) {
      P->pred[ci] += get_vlc(r, P->dct[ci], NULL);
      q[0] = (int16_t)(P->pred[ci] * (1 << P->al));
    } else if (get_bits(r, 1)) {
      q[0] = (int16_t)(q[0] | (1 << P->al));
    }
    return r->error;
  }
  const int p1 = 1 << P->al;
  int k = P->ss;
  if (P->ah == 0) {
    if (P->eobrun > 0) {
      P->eobrun--;
      return 0;
    }
    while (k <= P->se) {
      int v = get_vlc(r, P->act, &code);
      if (r->error) return r->error;
      int run = code >> 4, size = code & 15;
      if (size) {
        k += run;
        if (k > P->se) return ORC_SYNTAX_ERROR;
        q[kDeZigZag[k++]] = (int16_t)(v * p1);
      } else if (run == 15) {
        k += 16;
        if (k > P->se + 1) return ORC_SYNTAX_ERROR;
      } else {
        P->eobrun = (1 << run) - 1 + (run ? get_bits(r, run) : 0);
        break;
      }
    }
    return r->error;
  }
  /* refinement: the correction bit of a nonzero coefficient */
#define REFINE(x)                                                   \
  do {                                                              \
    if (get_bits(r, 1) && !((x) & p1)) (x) = (int16_t)((x) + ((x) >= 0 ? p1 : -p1)); \
  } while (0)
  if (P->eobrun == 0) {
    while (k <= P->se) {
      /* a size-1 symbol's one extra bit is the new coefficient's sign, which
       * EXTEND turns into +1 or -1 */
      const int v = get_vlc(r, P->act, &code);
      if (r->error) return r->error;
      const int run = code >> 4, size = code & 15;
      if (size > 1) return ORC_SYNTAX_ERROR;
      if (!size && run != 15) {
        P->eobrun = (1 << run) + (run ? get_bits(r, run) : 0);
        break;
      }
      int left = run; /* zero coefficients to pass before the target one */
      while (k <= P->se) {
        int16_t *x = &q[kDeZigZag[k]];
        if (*x) REFINE(*x);
        else if (left-- == 0) break;
        k++;
      }
      if (k > P->se) {
        if (size) return ORC_SYNTAX_ERROR;
        break;
      }
      if (size) q[kDeZigZag[k]] = (int16_t)(v * p1);
      k++;
    }
  }
  if (P->eobrun > 0) {
    for (; k <= P->se; k++) {
      int16_t *x = &q[kDeZigZag[k]];
      if (*x) REFINE(*x);
    }
    P->eobrun--;
  }
#undef REFINE
  return r->error;
}

static int decode_scans_prog(orc_jpeg *j, int16_t *qcoef, int32_t *work[3]) {
  const long bpm = j->blocks_per_mcu;
  int16_t *q = qcoef ? qcoef : (int16_t *)malloc(sizeof(int16_t) * 64 * (size_t)j->num_blocks);
  vlc_t *lut = (vlc_t *)malloc(sizeof(vlc_t) * 65536 * 4);
  if (!q || !lut) {
    if (!qcoef) free(q);
    free(lut);
    FAIL(ORC_OOM_ERROR);
  }
  memset(q, 0, sizeof(int16_t) * 64 * (size_t)j->num_blocks);
  int slot0[3] = {0, 0, 0};
  for (int c = 1; c < j->num_channels; c++) slot0[c] = slot0[c - 1] + j->ch[c - 1].samples_x * j->ch[c - 1].samples_y;
  int err = 0;
  for (int k = 0; k < j->num_scans && !err; k++) {
    const int ns = j->scans[k].ncomp;
    prog_t P;
    memset(&P, 0, sizeof(P));
    P.ss = j->scans[k].ss;
    P.se = j->scans[k].se;
    P.ah = j->scans[k].ah;
    P.al = j->scans[k].al;
    /* lut: DC tables 0 and 1, then the scan's AC table */
    build_lut(j->scans[k].bits[0], j->scans[k].vals[0], lut);
    build_lut(j->scans[k].bits[1], j->scans[k].vals[1], lut + 65536);
    build_lut(j->scans[k].bits[j->scans[k].ac_id], j->scans[k].vals[j->scans[k].ac_id], lut + 2 * 65536);
    for (int i = 0; i < ns; i++) P.dct[i] = lut + 65536 * j->scans[k].dcs[i];
    P.act = lut + 2 * 65536;
    uint8_t *ub = (uint8_t *)malloc(j->scans[k].scan_len + 2);
    if (!ub) { err = ORC_OOM_ERROR; break; }
    size_t ulen = orc_unstuff(j->scans[k].scan, j->scans[k].scan_len, ub);
    reader_t rd = {ub, ub, ub + ulen, 0, 0, 0};
    P.r = rd;
    const int ri = j->scans[k].restart_interval;
    int rcount = ri, next_rst = 0;
    long units, bw = 0;
    const orc_channel *c0 = &j->ch[j->scans[k].comps[0]];
    if (ns > 1) {
      units = (long)j->num_blocks_x * j->num_blocks_y;
    } else {
      bw = (c0->width + 7) / 8;
      units = bw * ((c0->height + 7) / 8);
    }
    for (long u = 0; u < units && !err; u++) {
      if (ns > 1) {
        for (int i = 0; i < ns && !err; i++) {
          const int c = j->scans[k].comps[i];
          const int n = j->ch[c].samples_x * j->ch[c].samples_y;
          for (int b = 0; b < n && !err; b++) err = prog_block(&P, i, q + (u * bpm + slot0[c] + b) * 64);
        }
      } else {
        const int c = j->scans[k].comps[0], h = c0->samples_x, v = c0->samples_y;
        const long by = u / bw, bx = u % bw;
        const long mcu = (by / v) * j->num_blocks_x + bx / h;
        err = prog_block(&P, 0, q + (mcu * bpm + slot0[c] + (by % v) * h + bx % h) * 64);
      }
      /* between restart intervals (src/decodeScanGPU.cu:476-486, per unit of
       * the scan): byte-align and read RSTn; after the scan's last interval
       * nothing is read (data left there is ignored) */
      if (!err && ri && !(--rcount) && u + 1 < units) {
        P.r.num_bufbits &= 0xF8;
        int m = get_bits(&P.r, 16);
        if ((m & 0xFFF8) != 0xFFD0 || (m & 7) != next_rst) err = ORC_SYNTAX_ERROR;
        next_rst = (next_rst + 1) & 7;
        rcount = ri;
        P.pred[0] = P.pred[1] = P.pred[2] = 0;
        P.eobrun = 0;
      }
    }
    free(ub);
  }
  free(lut);
  /* working space: each component's blocks in decode order, dequantised */
  for (int c = 0; c < j->num_channels && !err && work; c++) {
    if (!work[c]) continue;
    const int n = j->ch[c].samples_x * j->ch[c].samples_y;
    int32_t *w = work[c];
    memset(w, 0, sizeof(int32_t) * (size_t)j->ch[c].stride * j->ch[c].plane_height);
    for (long m = 0; m < (long)j->num_blocks_x * j->num_blocks_y; m++)
      for (int b = 0; b < n; b++, w += 64)
        for (int z = 0; z < 64; z++)
          w[kDeZigZag[z]] = q[(m * bpm + slot0[c] + b) * 64 + kDeZigZag[z]] * j->comp_dq16[c][z];
  }
  if (!qcoef) free(q);
  if (err) FAIL(err);
  return 0;
}

int orc_decode_scan(const orc_jpeg *jc, int16_t *qcoef, int32_t *work[3], uint32_t *bitpos) {
  orc_jpeg *j = (orc_jpeg *)jc; /* only error fields are written */
  if (!j->have_scan) FAIL(ORC_SYNTAX_ERROR);
  if (j->progressive) return decode_scans_prog(j, qcoef, work);
  if (j->num_scans) return decode_scans_ni(j, qcoef, work);
  vlc_t *lut = (vlc_t *)malloc(sizeof(vlc_t) * 65536 * 4);
  uint8_t *ub = (uint8_t *)malloc(j->scan_len + 2);
  if (!lut || !ub) { free(lut); free(ub); FAIL(ORC_OOM_ERROR); }
  for (int t = 0; t < 4; t++) build_lut(j->dht_bits[t], j->dht_vals[t], lut + 65536 * t);
  size_t ulen = orc_unstuff(j->scan, j->scan_len, ub);
  reader_t r = {ub, ub, ub + ulen, 0, 0, 0};
  int dc[3] = {0, 0, 0};
  int32_t *wpos[3] = {0, 0, 0};
  for (int c = 0; c < j->num_channels; c++) {
    if (work && work[c]) {
      wpos[c] = work[c];
      memset(work[c], 0, sizeof(int32_t) * (size_t)j->ch[c].stride * j->ch[c].plane_height);
    }
  }
  int ri = j->restart_interval, rcount = ri, next_rst = 0;
  long blk = 0;
  int err = 0;
  for (int by = 0; by < j->num_blocks_y && !err; by++)
    for (int bx = 0; bx < j->num_blocks_x && !err; bx++) {
      for (int c = 0; c < j->num_channels && !err; c++) {
        const orc_channel *ch = &j->ch[c];
        const vlc_t *dct = lut + 65536 * ch->dc_id, *act = lut + 65536 * ch->ac_id;
        const uint16_t *dq = j->dq16[ch->dq_id];
        for (int sy = 0; sy < ch->samples_y && !err; sy++)
          for (int sx = 0; sx < ch->samples_x && !err; sx++) {
            /* src/decodeScanGPU.cu:78-100 (decodeBlock) */
            if (bitpos) bitpos[blk] = (uint32_t)((r.pos - r.buf) * 8 - r.num_bufbits);
            int16_t q[64];
            memset(q, 0, sizeof(q));
            dc[c] += get_vlc(&r, dct, NULL);
            q[0] = (int16_t)dc[c];
            int32_t *w = wpos[c];
            if (w) { memset(w, 0, 64 * sizeof(int32_t)); w[0] = dc[c] * dq[0]; }
            int coef = 0;
            uint8_t code = 0;
            do {
              int value = get_vlc(&r, act, &code);
              if (!code) break;
              if (!(code & 0x0F) && code != 0xF0) { err = ORC_SYNTAX_ERROR; break; }
              coef += (code >> 4) + 1;
              if (coef > 63) { err = ORC_SYNTAX_ERROR; break; }
              q[kDeZigZag[coef]] = (int16_t)value;
              if (w) w[kDeZigZag[coef]] = value * dq[coef];
            } while (coef < 63);
            if (r.error) err = r.error;
            if (qcoef) memcpy(qcoef + blk * 64, q, sizeof(q));
            if (w) wpos[c] += 64;
            blk++;
          }
      }
      /* src/decodeScanGPU.cu:476-486 */
      if (!err && ri && !(--rcount) && r.pos < r.end) {
        r.num_bufbits &= 0xF8;
        int m = get_bits(&r, 16);
        if ((m & 0xFFF8) != 0xFFD0 || (m & 7) != next_rst) err = ORC_SYNTAX_ERROR;
        next_rst = (next_rst + 1) & 7;
        rcount = ri;
        dc[0] = dc[1] = dc[2] = 0;
      }
    }
  free(lut);
  free(ub);
  if (err) FAIL(err);
  return 0;
}

/* ------------------------------------------------------------------ iDCT -- */
static inline uint8_t clip8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

/* CPUonly/src/pixelTransformCPU.c:14-57 */
static void idct_row(int *D, int shortcut) {
  int x0, x1, x2, x3, x4, x5, x6, x7, x8;
  x1 = D[4] << 11; x2 = D[6]; x3 = D[2]; x4 = D[1]; x5 = D[7]; x6 = D[5]; x7 = D[3];
  if (shortcut && !(x1 | x2 | x3 | x4 | x5 | x6 | x7)) {
    D[0] = D[1] = D[2] = D[3] = D[4] = D[5] = D[6] = D[7] = D[0] << 3;
    return;
  }
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
  D[0] = (x7 + x1) >> 8; D[1] = (x3 + x2) >> 8; D[2] = (x0 + x4) >> 8; D[3] = (x8 + x6) >> 8;
  D[4] = (x8 - x6) >> 8; D[5] = (x0 - x4) >> 8; D[6] = (x3 - x2) >> 8; D[7] = (x7 - x1) >> 8;
}

/* CPUonly/src/pixelTransformCPU.c:60-107 */
static void idct_col(const int *D, uint8_t *out, int stride, int shortcut) {
  int x0, x1, x2, x3, x4, x5, x6, x7, x8;
  x1 = D[8 * 4] << 8; x2 = D[8 * 6]; x3 = D[8 * 2]; x4 = D[8 * 1];
  x5 = D[8 * 7]; x6 = D[8 * 5]; x7 = D[8 * 3];
  if (shortcut && !(x1 | x2 | x3 | x4 | x5 | x6 | x7)) {
    uint8_t v = clip8(((D[0] + 32) >> 6) + 128);
    for (int i = 0; i < 8; i++) out[i * stride] = v;
    return;
  }
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
  out[0 * stride] = clip8(((x7 + x1) >> 14) + 128);
  out[1 * stride] = clip8(((x3 + x2) >> 14) + 128);
  out[2 * stride] = clip8(((x0 + x4) >> 14) + 128);
  out[3 * stride] = clip8(((x8 + x6) >> 14) + 128);
  out[4 * stride] = clip8(((x8 - x6) >> 14) + 128);
  out[5 * stride] = clip8(((x0 - x4) >> 14) + 128);
  out[6 * stride] = clip8(((x3 - x2) >> 14) + 128);
  out[7 * stride] = clip8(((x7 - x1) >> 14) + 128);
}

static void idct_int_impl(const int32_t in[64], uint8_t *out, int stride, int shortcut) {
  int D[64];
  for (int i = 0; i < 64; i++) D[i] = in[i];
  for (int r = 0; r < 64; r += 8) idct_row(D + r, shortcut);
  for (int c = 0; c < 8; c++) idct_col(D + c, out + c, stride, shortcut);
}
void orc_idct_int(const int32_t in[64], uint8_t *out, int stride) { idct_int_impl(in, out, stride, 1); }
void orc_idct_int_noshortcut(const int32_t in[64], uint8_t *out, int stride) {
  idct_int_impl(in, out, stride, 0);
}

/* T.81 A.3.3: s(y,x) = 1/4 sum_v sum_u C(u)C(v) S(v,u) cos((2x+1)u pi/16) cos((2y+1)v pi/16) */
static double g_c[8][8];
static int g_c_ready = 0;
void orc_idct_f64_raw(const int32_t in[64], double out[64]) {
  if (!g_c_ready) {
    for (int u = 0; u < 8; u++)
      for (int x = 0; x < 8; x++)
        g_c[u][x] = (u == 0 ? sqrt(0.125) : 0.5) * cos((2 * x + 1) * u * M_PI / 16.0);
    g_c_ready = 1;
  }
  double t[64];
  for (int v = 0; v < 8; v++)
    for (int x = 0; x < 8; x++) {
      double s = 0;
      for (int u = 0; u < 8; u++) s += g_c[u][x] * in[v * 8 + u];
      t[v * 8 + x] = s;
    }
  for (int y = 0; y < 8; y++)
    for (int x = 0; x < 8; x++) {
      double s = 0;
      for (int v = 0; v < 8; v++) s += g_c[v][y] * t[v * 8 + x];
      out[y * 8 + x] = s + 128.0;
    }
}
void orc_idct_f64(const int32_t in[64], uint8_t *out, int stride) {
  double o[64];
  orc_idct_f64_raw(in, o);
  for (int y = 0; y < 8; y++)
    for (int x = 0; x < 8; x++) out[y * stride + x] = clip8((int)floor(o[y * 8 + x] + 0.5));
}

/* src/pixelTransformGPU.cu:120-131 block placement; decode order is MCU raster,
 * then component, then sample row, then sample column (src/decodeScanCPU.cu:118-137). */
int orc_planes(const orc_jpeg *j, const int16_t *qcoef, uint8_t *planes[3], int mode) {
  long blk = 0;
  for (int by = 0; by < j->num_blocks_y; by++)
    for (int bx = 0; bx < j->num_blocks_x; bx++)
      for (int c = 0; c < j->num_channels; c++) {
        const orc_channel *ch = &j->ch[c];
        const uint16_t *dq = j->num_scans || j->progressive ? j->comp_dq16[c] : j->dq16[ch->dq_id];
        for (int sy = 0; sy < ch->samples_y; sy++)
          for (int sx = 0; sx < ch->samples_x; sx++, blk++) {
            int32_t deq[64];
            for (int k = 0; k < 64; k++) deq[kDeZigZag[k]] = qcoef[blk * 64 + kDeZigZag[k]] * dq[k];
            size_t off = ((size_t)(by * ch->samples_y + sy) * ch->stride + (size_t)(bx * ch->samples_x + sx)) << 3;
            if (mode == 0) orc_idct_int(deq, planes[c] + off, ch->stride);
            else orc_idct_f64(deq, planes[c] + off, ch->stride);
          }
      }
  return 0;
}

/* CPUonly/src/pixelTransformCPU.c:110-172 */
int orc_colour(const orc_jpeg *j, uint8_t *const planes[3], uint8_t *out) {
  const int W = j->width, H = j->height;
  if (j->num_channels == 1) {
    const orc_channel *c = &j->ch[0];
    for (int y = 0; y < H; y++) memcpy(out + (size_t)y * W, planes[0] + (size_t)y * c->stride, (size_t)W);
    return 0;
  }
  int xs[3], ys[3];
  for (int c = 0; c < 3; c++) {
    int w = j->ch[c].width, h = j->ch[c].height;
    xs[c] = ys[c] = 0;
    while (w < W) { w <<= 1; xs[c]++; }
    while (h < H) { h <<= 1; ys[c]++; }
  }
  for (int y = 0; y < H; y++) {
    const uint8_t *py = planes[0] + (size_t)(y >> ys[0]) * j->ch[0].stride;
    const uint8_t *pb = planes[1] + (size_t)(y >> ys[1]) * j->ch[1].stride;
    const uint8_t *pr = planes[2] + (size_t)(y >> ys[2]) * j->ch[2].stride;
    uint8_t *o = out + (size_t)y * W * 3;
    for (int x = 0; x < W; x++) {
      int Y = py[x >> xs[0]] << 8, cb = pb[x >> xs[1]] - 128, cr = pr[x >> xs[2]] - 128;
      o[3 * x + 0] = clip8((Y + 359 * cr + 128) >> 8);
      o[3 * x + 1] = clip8((Y - 88 * cb - 183 * cr + 128) >> 8);
      o[3 * x + 2] = clip8((Y + 454 * cb + 128) >> 8);
    }
  }
  return 0;
}

int orc_decode_file(const uint8_t *file, size_t size, int idct_mode, uint8_t *out, size_t out_cap,
                    int *width, int *height, int *ncomp) {
  return orc_decode_file_ext(file, size, 0, idct_mode, out, out_cap, width, height, ncomp);
}

int orc_decode_file_ext(const uint8_t *file, size_t size, int ext, int idct_mode, uint8_t *out, size_t out_cap,
                        int *width, int *height, int *ncomp) {
  orc_jpeg j;
  int e = orc_parse_ext(file, size, ext, &j);
  if (e) return e;
  *width = j.width; *height = j.height; *ncomp = j.num_channels;
  if ((size_t)j.width * j.height * j.num_channels > out_cap) return ORC_OOM_ERROR;
  int16_t *q = (int16_t *)malloc(sizeof(int16_t) * 64 * (size_t)j.num_blocks);
  uint8_t *pl[3] = {0, 0, 0};
  if (!q) return ORC_OOM_ERROR;
  e = orc_decode_scan(&j, q, NULL, NULL);
  for (int c = 0; c < j.num_channels && !e; c++) {
    pl[c] = (uint8_t *)malloc((size_t)j.ch[c].stride * j.ch[c].plane_height);
    if (!pl[c]) e = ORC_OOM_ERROR;
  }
  if (!e) e = orc_planes(&j, q, pl, idct_mode);
  if (!e) e = orc_colour(&j, pl, out);
  for (int c = 0; c < 3; c++) free(pl[c]);
  free(q);
  return e;
}
