/* Harness around the reference src/decodeScanCPU.cu (test infrastructure).
 * Fills the reference JPGReader from the oracle's parse of the file and runs
 * the reference decodeScanCPU on it, returning the dequantised working_space
 * per component.  The 65536-entry Huffman LUTs are built with the spreading
 * rule of src/format.cu:371-391 (format.cu itself needs CUDA and is not built). */
#include <stdlib.h>
#include <string.h>

#include <format.h>          /* reference include/format.h */
#include <decodeScanCPU.h>   /* reference include/decodeScanCPU.h */
#include "jpeg_oracle.h"

unsigned short read16(const unsigned char *pos) { return (unsigned short)((pos[0] << 8) | pos[1]); }

static void fill_lut(const uint8_t bits[16], const uint8_t *vals, DhtVlc *vlc) {
  int remain = 65536, spread = 65536, k = 0;
  for (int l = 1; l <= 16; l++) {
    spread >>= 1;
    int count = bits[l - 1];
    remain -= count << (16 - l);
    for (int i = 0; i < count; i++, k++)
      for (int s = spread; s; s--, vlc++) { vlc->num_bits = (unsigned char)l; vlc->tuple = vals[k]; }
  }
  while (remain-- > 0) { vlc->num_bits = 0; vlc->tuple = 0; vlc++; }
}

extern "C" int ref_scan_decode(const unsigned char *file, long size, int32_t *out[3]) {
  orc_jpeg oj;
  int e = orc_parse(file, (size_t)size, &oj);
  if (e) return e;
  JPGReader *r = (JPGReader *)calloc(1, sizeof(JPGReader));
  unsigned char *buf = (unsigned char *)malloc((size_t)size);
  DhtVlc *luts = (DhtVlc *)malloc(sizeof(DhtVlc) * 65536 * 4);
  memcpy(buf, file, (size_t)size);
  for (int t = 0; t < 4; t++) {
    fill_lut(oj.dht_bits[t], oj.dht_vals[t], luts + 65536 * t);
    r->vlc_tables[t] = luts + 65536 * t;
  }
  r->width = (unsigned short)oj.width; r->height = (unsigned short)oj.height;
  r->num_blocks_x = (unsigned short)oj.num_blocks_x; r->num_blocks_y = (unsigned short)oj.num_blocks_y;
  r->block_size_x = (unsigned short)oj.block_size_x; r->block_size_y = (unsigned short)oj.block_size_y;
  r->num_channels = (unsigned char)oj.num_channels;
  r->restart_interval = oj.restart_interval;
  memcpy(r->dq_tables, oj.dq, sizeof(oj.dq));
  for (int c = 0; c < oj.num_channels; c++) {
    ColourChannel *ch = &r->channels[c];
    ch->id = oj.ch[c].id; ch->dq_id = oj.ch[c].dq_id;
    ch->samples_x = oj.ch[c].samples_x; ch->samples_y = oj.ch[c].samples_y;
    ch->width = oj.ch[c].width; ch->height = oj.ch[c].height; ch->stride = oj.ch[c].stride;
    ch->dc_cumulative_val = 0;
    size_t n = (size_t)oj.ch[c].stride * oj.ch[c].plane_height;
    ch->working_space.mem = (int *)calloc(n, sizeof(int));
    ch->working_space.size = ch->working_space.max_size = (unsigned)n;
    ch->working_space_pos = ch->working_space.mem;
  }
  /* src/format.cu:166-167: end excludes the EOI; pos at the SOS length field */
  r->end = buf + size - 2;
  r->pos = buf + (oj.scan - file);
  /* walk back over the SOS header: its length field sits 2+len before the data */
  {
    const unsigned char *p = buf + 2;
    while (p < r->end) {
      unsigned m = p[1];
      p += 2;
      if (m == 0xDA) { r->pos = (unsigned char *)p; break; }
      if (m == 0xD8 || m == 0xD9) continue;
      p += read16(p);
    }
  }
  decodeScanCPU(r);
  e = r->error;
  for (int c = 0; c < oj.num_channels; c++) {
    if (out && out[c]) memcpy(out[c], r->channels[c].working_space.mem, r->channels[c].working_space.size * sizeof(int));
    free(r->channels[c].working_space.mem);
  }
  free(luts); free(buf); free(r);
  return e;
}
