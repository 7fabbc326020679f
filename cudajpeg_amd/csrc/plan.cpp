// Host planning of a decode batch (see plan.h).
//
// Host side of what the reference does in decodeScanGPU before its kernels
// (src/decodeScanGPU.cu:104-151: table reshuffle, unstuffing, scan upload),
// generalised to many images and restart segments.
#include "plan.h"

#include <string.h>

#include <algorithm>

#include "parse_internal.h"

namespace cjd {

namespace {

// zigzag index -> natural index (include/format.h:117-120)
const uint8_t kZZ[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15,
                         23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman decode tables (T.81 C.2 / F.2.2.3).  Codes are assigned
// in the same order as the reference's LUT spreading (src/format.cu:371-391).
}  // namespace

void build_table(const uint8_t bits[16], const uint8_t vals[256], HuffTable& t) {
  memset(&t, 0, sizeof(t));
  int code = 0, k = 0;
  for (int l = 1; l <= 16; ++l) {
    const int n = bits[l - 1];
    t.valoff[l] = k - code;
    if (l <= kLutBits) {
      for (int i = 0; i < n; ++i) {
        const int lo = (code + i) << (kLutBits - l), hi = (code + i + 1) << (kLutBits - l);
        for (int e = lo; e < hi && e < kLutSize; ++e) t.lut[e] = (uint32_t)l | ((uint32_t)vals[k + i] << 8);
      }
    }
    code += n;
    k += n;
    t.lim[l] = (uint32_t)code << (16 - l);
    code <<= 1;
  }
  memcpy(t.vals, vals, 256);
}

namespace {

// src/decodeScanGPU.cu:11-24 (remove 0x00 after 0xFF, keep RSTn, drop other
// marker bytes) with the scan split at RSTn when a restart interval is set.
// Returns false on an out-of-sequence RSTn (src/decodeScanGPU.cu:480-482).
bool unstuff_segments(const uint8_t* src, size_t len, bool split, std::vector<std::vector<uint8_t>>& segs) {
  segs.clear();
  segs.emplace_back();
  segs.back().reserve(len);
  int next_rst = 0;
  size_t s = 0;
  while (s < len) {
    const uint8_t b = src[s++];
    if (b != 0xFF) {
      segs.back().push_back(b);
      continue;
    }
    const uint8_t nx = s < len ? src[s] : 0xFF;
    ++s;
    if (nx == 0x00) {
      segs.back().push_back(0xFF);
      continue;
    }
    if ((nx & 0xF8) == 0xD0) {
      if (split) {
        if ((nx & 7) != next_rst) return false;
        next_rst = (next_rst + 1) & 7;
        segs.emplace_back();
        continue;
      }
      segs.back().push_back(0xFF);
      segs.back().push_back(nx);
      continue;
    }
    segs.back().push_back(0xFF);
  }
  return true;
}

}  // namespace

static int stage_image(Plan* b, ImageRec& r, const uint8_t* file) {
  const cjd_frame_info& f = r.f;
  if (f.blocks_per_mcu > kMaxSlots) return CJD_UNSUPPORTED_ERROR;
  std::vector<std::vector<uint8_t>> parts;
  if (!unstuff_segments(file + f.scan_offset, (size_t)f.scan_length, f.restart_interval > 0, parts))
    return CJD_SYNTAX_ERROR;
  const int64_t nmcu = (int64_t)f.mcus_x * f.mcus_y;
  int64_t nseg_expected = 1;
  if (f.restart_interval > 0) nseg_expected = (nmcu + f.restart_interval - 1) / f.restart_interval;
  // a trailing RSTn after the last interval leaves an empty extra segment
  while ((int64_t)parts.size() > nseg_expected && parts.back().empty()) parts.pop_back();
  if ((int64_t)parts.size() != nseg_expected) return CJD_SYNTAX_ERROR;

  ImageDesc d{};
  d.width = f.width;
  d.height = f.height;
  d.ncomp = f.num_channels;
  d.mcus_x = f.mcus_x;
  d.mcus_y = f.mcus_y;
  d.nslots = f.blocks_per_mcu;
  d.mcu_w = f.mcu_width;
  d.mcu_h = f.mcu_height;
  d.table_base = (int32_t)b->tables.size();
  d.qtab_base = (int32_t)(b->qtabs.size() / 64);
  int slot = 0;
  for (int c = 0; c < f.num_channels; ++c) {
    d.comp_first_slot[c] = slot;
    d.comp_dq[c] = (uint8_t)f.dq_id[c];
    d.comp_sx[c] = (uint8_t)f.samples_x[c];
    d.comp_sy[c] = (uint8_t)f.samples_y[c];
    int w = f.comp_width[c], h = f.comp_height[c], xs = 0, ys = 0;
    while (w < f.width) { w <<= 1; ++xs; }   // CPUonly/src/pixelTransformCPU.c:113-114
    while (h < f.height) { h <<= 1; ++ys; }
    d.comp_xshift[c] = (uint8_t)xs;
    d.comp_yshift[c] = (uint8_t)ys;
    d.comp_stride[c] = f.comp_stride[c];
    d.comp_plane_h[c] = f.comp_plane_height[c];
    for (int k = 0; k < f.samples_x[c] * f.samples_y[c]; ++k, ++slot) {
      d.slot_comp[slot] = (uint8_t)c;
      d.slot_dc[slot] = (uint8_t)f.dc_id[c];
      d.slot_ac[slot] = (uint8_t)f.ac_id[c];
    }
    int t = 0;
    while (t < d.ntab && d.tab_dq[t] != d.comp_dq[c]) ++t;
    if (t == d.ntab) d.tab_dq[d.ntab++] = d.comp_dq[c];
    d.comp_tix[c] = (uint8_t)t;
  }
  for (int t = 0; t < 4; ++t) {
    HuffTable ht;
    build_table(f.huff_bits[t], f.huff_vals[t], ht);
    b->tables.push_back(ht);
    for (int k = 0; k < 64; ++k) b->qtabs.push_back(0);
    uint16_t* q = &b->qtabs[b->qtabs.size() - 64];
    for (int k = 0; k < 64; ++k) q[kZZ[k]] = f.dq_tables[t][k];
  }
  d.coef_base = b->total_blocks;
  d.nblocks = f.num_blocks;
  r.coef_base = d.coef_base;
  b->total_blocks += f.num_blocks;
  const uint64_t ob = (uint64_t)f.width * f.height * (f.num_channels == 3 ? 3 : 1);
  r.out_offset = b->out_bytes;
  r.out_bytes = ob;
  d.out_offset = r.out_offset;
  b->out_bytes += (ob + 255) & ~255ull;
  for (int c = 0; c < 3; ++c) d.plane_offset[c] = ~0ull;
  if (b->want_planes) {
    for (int c = 0; c < f.num_channels; ++c) {
      const uint64_t pb = (uint64_t)f.comp_stride[c] * f.comp_plane_height[c];
      d.plane_offset[c] = r.plane_offset[c] = b->out_bytes;
      b->out_bytes += (pb + 255) & ~255ull;
    }
  }
  d.error_index = (int32_t)(&r - &b->recs[0]);
  r.desc = (uint32_t)b->images.size();
  b->images.push_back(d);
  r.pixels = (int64_t)f.width * f.height;
  b->total_pixels += r.pixels;

  // segments, tiles and entropy work items
  const int nh = f.blocks_per_mcu;
  // one lane per tile; an item after the first also decodes its lead-in tile
  const uint32_t per_first = b->item_tiles, per_next = b->item_tiles - 1;
  r.seg_first = (uint32_t)b->segs.size();
  for (size_t s = 0; s < parts.size(); ++s) {
    const std::vector<uint8_t>& data = parts[s];
    SegmentDesc sd{};
    sd.word_base = b->words.size();
    sd.nbits = (uint32_t)(data.size() * 8);
    // 16-byte aligned start (the decoder reads 4-word chunks) and
    // kSegPadWords words of 0xFF padding past the data
    b->words.resize((b->words.size() + 3) & ~(size_t)3, 0xFFFFFFFFu);
    sd.word_base = b->words.size();
    const size_t nw = (((data.size() + 3) / 4 + kSegPadWords) + 3) & ~(size_t)3;
    sd.nwords = (uint32_t)nw;
    const size_t w0 = b->words.size();
    b->words.resize(w0 + nw, 0xFFFFFFFFu);
    memcpy(&b->words[w0], data.data(), data.size());
    b->scan_bytes += (int64_t)data.size();
    sd.image = r.desc;
    sd.tile_bits = b->tile_bits;
    // sync points: trajectories from the slot hypotheses mostly agree within
    // a few blocks, and a main trajectory that starts off the true one still
    // joins it before the tile end, so a short first sync point suffices
    sd.sync_bits = b->sync_bits ? b->sync_bits : std::max<uint32_t>(64, b->tile_bits / 32);
    sd.sync1_bits = b->sync1_bits ? b->sync1_bits : sd.sync_bits;
    if (sd.sync1_bits > sd.sync_bits) sd.sync1_bits = sd.sync_bits;
    sd.ntiles = std::max<uint32_t>(1, (sd.nbits + b->tile_bits - 1) / b->tile_bits);
    const int64_t mcu0 = (int64_t)s * (f.restart_interval > 0 ? f.restart_interval : nmcu);
    const int64_t mcus = std::min<int64_t>(nmcu - mcu0, f.restart_interval > 0 ? f.restart_interval : nmcu);
    sd.nblocks = (uint32_t)(mcus * nh);
    sd.first_block = d.coef_base + mcu0 * nh;
    sd.first_item = (uint32_t)b->items.size();
    uint32_t t0 = 0, prev = ~0u;
    while (t0 < sd.ntiles) {
      WorkItem wi{};
      wi.segment = (uint32_t)b->segs.size();
      wi.t0 = t0;
      wi.ntiles = std::min(sd.ntiles - t0, t0 == 0 ? per_first : per_next);
      wi.prev_item = prev;
      prev = (uint32_t)b->items.size();
      b->items.push_back(wi);
      t0 += wi.ntiles;
    }
    b->segs.push_back(sd);
  }
  r.seg_count = (uint32_t)b->segs.size() - r.seg_first;

  // pixel work items: runs of MCUs within one MCU row
  const int G = std::max(1, kPixBlocks / nh);
  r.pix_first = (uint32_t)b->pix.size();
  for (int my = 0; my < f.mcus_y; ++my)
    for (int mx = 0; mx < f.mcus_x; mx += G) {
      PixItem pi{};
      pi.image = r.desc;
      pi.mcu_first = (uint32_t)(my * f.mcus_x + mx);
      pi.count = (uint32_t)std::min(G, f.mcus_x - mx);
      b->pix.push_back(pi);
    }
  r.pix_count = (uint32_t)b->pix.size() - r.pix_first;
  return 0;
}

void plan_batch(Plan& b, int n, const uint8_t* const* files, const size_t* sizes, int* errors) {
  b.recs.assign((size_t)n, ImageRec{});
  b.images.clear();
  b.segs.clear();
  b.items.clear();
  b.pix.clear();
  b.tables.clear();
  b.qtabs.clear();
  b.words.clear();
  b.total_blocks = b.total_pixels = b.scan_bytes = 0;
  b.out_bytes = 0;
  for (int i = 0; i < n; ++i) {
    ImageRec& r = b.recs[(size_t)i];
    ParseState st;
    r.error = parse_until_scan(st, files[i], sizes[i], &r.f);
    if (!r.error && (!st.at_scan || !st.have_sof)) r.error = CJD_SYNTAX_ERROR;
    if (!r.error) r.error = stage_image(&b, r, files[i]);
    if (errors) errors[i] = r.error;
  }
}

}  // namespace cjd
