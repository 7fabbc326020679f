// Host planning of a decode batch (see plan.h).
//
// Host side of what the reference does in decodeScanGPU before its kernels
// (src/decodeScanGPU.cu:104-151: table reshuffle, unstuffing, scan upload),
// generalised to many images and restart segments.
#include "plan.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include "entropy_core.h"
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

void build_lut(const uint8_t bits[16], const uint8_t vals[256], bool ac, uint16_t* lut, uint16_t* l1, int& pool_used) {
  for (int e = 0; e < kL1Size; ++e) l1[e] = 0;
  int code = 0, k = 0;
  for (int l = 1; l <= 16; ++l) {
    const int n = bits[l - 1];
    for (int i = 0; i < n; ++i, ++k) {
      const int cd = code + i;
      const uint16_t ent = lut_entry(l, vals[k], ac);
      if (l <= kL1Bits) {
        const int lo = cd << (kL1Bits - l), hi = (cd + 1) << (kL1Bits - l);
        for (int e = lo; e < hi && e < kL1Size; ++e) l1[e] = ent;
        continue;
      }
      const int pre = cd >> (l - kL1Bits);
      if (pre >= kL1Size) continue;  // not a prefix code: no stream can reach it
      if (!(l1[pre] & 0x8000u)) {
        if (pool_used + kL2Size > kL2Pool) {
          l1[pre] = kLutSlow;
        } else {
          l1[pre] = (uint16_t)(0x8000u | (uint32_t)(4 * kL1Size + pool_used));
          for (int e = 0; e < kL2Size; ++e) lut[4 * kL1Size + pool_used + e] = 0;
          pool_used += kL2Size;
        }
      }
      if (l1[pre] == kLutSlow) continue;
      uint16_t* l2 = lut + (l1[pre] & 0x7FFFu);
      const int sub = (cd << (kL1Bits + kL2Bits - l)) & (kL2Size - 1);
      for (int e = sub; e < sub + (1 << (kL1Bits + kL2Bits - l)); ++e) l2[e] = ent;
    }
    code = (code + n) << 1;
  }
}

// src/decodeScanGPU.cu:11-24 (remove 0x00 after 0xFF, keep RSTn, drop other
// marker bytes) with the scan split at RSTn when a restart interval is set.
// Runs between 0xFF bytes are found with memchr and block-copied.  With
// dst == nullptr only the segment bounds (in unstuffed bytes, segment k =
// [bounds[k], bounds[k+1])) are computed; otherwise segment k is written to
// dst + dst_off[k].  With chunks, the starting state of every chunk of
// kUnstuffChunk stuffed bytes is appended for the device unstuffer (src
// relative to the scan, seg relative to the image's first segment).  Returns
// false on an out-of-sequence RSTn (src/decodeScanGPU.cu:480-482).
bool unstuff_segments(const uint8_t* src, size_t len, bool split, std::vector<size_t>* bounds, uint8_t* dst,
                      const size_t* dst_off, std::vector<UnstuffChunk>* chunks) {
  if (bounds) bounds->assign(1, 0);
  if (chunks) chunks->clear();
  size_t o = 0, s = 0, seg = 0;
  uint8_t* out = dst ? dst + dst_off[0] : nullptr;
  int next_rst = 0;
  size_t q = 0;  // start of the next chunk
  // chunk at q: its first output byte is byte dst_at of segment seg
  auto chunk = [&](size_t dst_at, uint32_t skip) {
    UnstuffChunk c{};
    c.src = q;
    c.len = (uint32_t)std::min<size_t>(kUnstuffChunk, len - q);
    c.seg = (uint32_t)seg;
    c.dst = (uint32_t)dst_at;
    c.flags = skip | (split ? kChunkSplit : 0u) | (q + c.len == len ? kChunkScanEnd : 0u);
    chunks->push_back(c);
    q += kUnstuffChunk;
  };
  while (s < len) {
    const void* ff = memchr(src + s, 0xFF, len - s);
    const size_t e = ff ? (size_t)(static_cast<const uint8_t*>(ff) - src) : len;
    if (chunks)
      while (q < e) chunk(o + (q - s), 0);
    if (out) memcpy(out + o, src + s, e - s);
    o += e - s;
    s = e;
    if (s >= len) break;
    const size_t p = s;  // the pair (p, p + 1)
    if (chunks && q == p) chunk(o, 0);
    const uint8_t nx = p + 1 < len ? src[p + 1] : 0xFF;
    s = p + 2;
    if (nx == 0x00) {
      if (out) out[o] = 0xFF;
      ++o;
    } else if ((nx & 0xF8) == 0xD0 && split) {
      if ((nx & 7) != next_rst) return false;
      next_rst = (next_rst + 1) & 7;
      if (bounds) bounds->push_back(bounds->back() + o);
      ++seg;
      if (dst) out = dst + dst_off[seg];
      o = 0;
    } else if ((nx & 0xF8) == 0xD0) {
      if (out) {
        out[o] = 0xFF;
        out[o + 1] = nx;
      }
      o += 2;
    } else {
      if (out) out[o] = 0xFF;
      ++o;
    }
    // a chunk that starts at the pair's second byte
    if (chunks && q == p + 1 && p + 1 < len) chunk(o, kChunkSkipFirst);
  }
  if (bounds) bounds->push_back(bounds->back() + o);
  return true;
}

namespace {

// One scan of an image: the interleaved scan, or (CJD_EXT_NONINTERLEAVED) a
// one-component scan with the tables and restart interval in effect at it.
struct ScanWork {
  int comp = -1;                  // -1: the interleaved scan; else the scan's (first) component
  int prog = -1;                  // a progressive scan: its index in ParseState::scans
  uint64_t off = 0, len = 0;      // its stuffed bytes in the file
  int ri = 0;                     // restart interval (MCUs of the scan: blocks for one component)
  int64_t units = 0;              // MCUs of the scan
  int h = 1, v = 1, bw = 0, slot0 = 0, dc_id = 0, ac_id = 2;  // one-component scan geometry and tables
  std::vector<size_t> bounds;     // unstuffed segment bounds
  std::vector<UnstuffChunk> chunks;  // device unstuffing: the scan's chunks (scan-relative)
  int64_t nseg = 0, seg_rel = 0, raw_rel = 0, chunk_rel = 0;  // offsets within the image's share
  const uint8_t (*bits)[16] = nullptr;  // its Huffman tables (BITS, HUFFVAL by table id)
  const uint8_t (*vals)[256] = nullptr;
};

// Sizes of one image's share of the plan (pass 1) and its offsets in the
// batch (pass 2).
struct ImageWork {
  std::vector<ScanWork> scans;
  ParseState st;                  // the parse (the tables of one-component scans)
  std::vector<size_t> bounds;     // unstuffed segment bounds of all scans, scan after scan
  int64_t nraw = 0, nchunks = 0;  // device unstuffing: the image's raw bytes and chunks
  int64_t raw0 = 0, chunk0 = 0;  // device unstuffing: offsets in the raw buffer and the chunk list
  int64_t tset0 = 0;             // first Huffman table set (one per scan)
  std::vector<int> level;         // progressive: each scan's level
  std::vector<int64_t> lev_count, lev_off;  // progressive: tasks per level, first task of each level
  uint32_t S = 4096, C = 512, items = kWG;  // entropy parameters for this image
  int64_t nseg = 0, nitems = 0, npix = 0, nwords = 0, nent = 0;
  int64_t seg0 = 0, item0 = 0, pix0 = 0, word0 = 0, desc = 0, coef0 = 0, ent0 = 0;
  uint64_t out0 = 0, plane0[3] = {~0ull, ~0ull, ~0ull};
};

// Words of one segment: 16-byte aligned start (the decoder reads 4-word
// chunks), the data, and kSegPadWords words of 0xFF padding.
inline int64_t seg_words(size_t dlen) { return (int64_t)((((dlen + 3) / 4 + kSegPadWords) + 3) & ~(size_t)3); }

inline uint32_t ntiles_of(size_t dlen, uint32_t S) {
  return std::max<uint32_t>(1, (uint32_t)((dlen * 8 + S - 1) / S));
}

// Stream and list sizes of the tiled decoder: a speculative or prefix stream
// holds S/2 + 128 symbols (a stream under two bits per symbol on average over
// a tile runs into the limit: the speculative pass then leaves the tile to
// its link, a link hands the image to the serial decoder), a DC list one
// block per two symbols (every valid block has a DC symbol and at least one
// AC symbol) plus slack.
struct TileGeom {
  uint32_t rs, rd, nck;
};
inline TileGeom tile_geom(uint32_t S, uint32_t C) {
  TileGeom g;
  g.rs = S / 2 + 128;
  g.rd = ((g.rs / 2 + 64) + 3) & ~3u;
  g.nck = (S - 1) / C;
  return g;
}

// Work items of a segment of nt tiles: the first takes per_first tiles, the
// others per_next each plus the lead-in; f(t0, ntiles) for each.
template <class F>
inline void for_items(uint32_t nt, uint32_t per_first, uint32_t per_next, F&& f) {
  uint32_t t0 = 0;
  while (t0 < nt) {
    const uint32_t n = std::min(nt - t0, t0 == 0 ? per_first : per_next);
    f(t0, n);
    t0 += n;
  }
}

// Pass 1 (per image, in parallel): parse, find the segments, count.
int image_pass1(const Plan& b, ImageRec& r, ImageWork& w, const uint8_t* file, size_t size) {
  ParseState& st = w.st;
  int e = parse_until_scan(st, file, size, &r.f, b.ext);
  if (!e && (!st.at_scan || !st.have_sof)) e = CJD_SYNTAX_ERROR;
  if (e) return e;
  const cjd_frame_info& f = r.f;
  if (f.blocks_per_mcu > kMaxSlots) return CJD_UNSUPPORTED_ERROR;
  w.scans.clear();
  w.level.clear();
  w.lev_count.clear();
  if (st.progressive) {
    // T.81 G.1.2: a scan of several components runs over the frame's MCUs,
    // a one-component scan over the component's own block grid; a scan's
    // level is one more than the last earlier scan's that codes one of its
    // coefficients
    for (int k = 0; k < st.nscans; ++k) {
      const ScanInfo& sc = st.scans[(size_t)k];
      ScanWork sw;
      sw.prog = k;
      sw.comp = sc.comp;
      sw.off = sc.offset;
      sw.len = sc.length;
      sw.ri = sc.restart_interval;
      if (sc.ncomp > 1) {
        sw.units = (int64_t)f.mcus_x * f.mcus_y;
      } else {
        sw.h = f.samples_x[sc.comp];
        sw.v = f.samples_y[sc.comp];
        sw.bw = (f.comp_width[sc.comp] + 7) / 8;
        sw.units = (int64_t)sw.bw * ((f.comp_height[sc.comp] + 7) / 8);
      }
      sw.bits = sc.huff_bits;
      sw.vals = sc.huff_vals;
      int lev = 0;
      for (int t = 0; t < k; ++t) {
        const ScanInfo& o = st.scans[(size_t)t];
        bool shared = false;
        for (int i = 0; i < sc.ncomp; ++i)
          for (int q = 0; q < o.ncomp; ++q) shared |= sc.comps[i] == o.comps[q];
        if (shared && o.ss <= sc.se && sc.ss <= o.se) lev = std::max(lev, w.level[(size_t)t] + 1);
      }
      w.level.push_back(lev);
      w.scans.push_back(std::move(sw));
    }
  } else if (!st.nscans) {
    ScanWork sw;
    sw.off = f.scan_offset;
    sw.len = f.scan_length;
    sw.ri = f.restart_interval;
    sw.units = (int64_t)f.mcus_x * f.mcus_y;
    sw.bits = f.huff_bits;
    sw.vals = f.huff_vals;
    w.scans.push_back(std::move(sw));
  } else {
    // one-component scans (T.81 A.2.2): the component's own block grid
    for (int k = 0; k < st.nscans; ++k) {
      const ScanInfo& sc = st.scans[k];
      ScanWork sw;
      sw.comp = sc.comp;
      sw.off = sc.offset;
      sw.len = sc.length;
      sw.ri = sc.restart_interval;
      sw.h = f.samples_x[sc.comp];
      sw.v = f.samples_y[sc.comp];
      sw.bw = (f.comp_width[sc.comp] + 7) / 8;
      sw.units = (int64_t)sw.bw * ((f.comp_height[sc.comp] + 7) / 8);
      for (int c = 0; c < sc.comp; ++c) sw.slot0 += f.samples_x[c] * f.samples_y[c];
      sw.dc_id = sc.dc_id;
      sw.ac_id = sc.ac_id;
      sw.bits = sc.huff_bits;
      sw.vals = sc.huff_vals;
      w.scans.push_back(std::move(sw));
    }
  }
  w.bounds.assign(1, 0);
  w.nseg = w.nraw = w.nchunks = 0;
  uint64_t scan_bytes = 0;
  int64_t coded = 0;
  for (ScanWork& sw : w.scans) {
    if (!unstuff_segments(file + sw.off, (size_t)sw.len, sw.ri > 0, &sw.bounds, nullptr, nullptr,
                          b.unstuff_device ? &sw.chunks : nullptr))
      return CJD_SYNTAX_ERROR;
    int64_t nseg_expected = 1;
    if (sw.ri > 0) nseg_expected = (sw.units + sw.ri - 1) / sw.ri;
    std::vector<size_t>& bd = sw.bounds;
    // a trailing RSTn after the last interval leaves an empty extra segment
    while ((int64_t)bd.size() - 1 > nseg_expected && bd[bd.size() - 1] == bd[bd.size() - 2]) bd.pop_back();
    if ((int64_t)bd.size() - 1 != nseg_expected) return CJD_SYNTAX_ERROR;
    sw.nseg = nseg_expected;
    sw.seg_rel = w.nseg;
    sw.raw_rel = w.nraw;
    sw.chunk_rel = w.nchunks;
    w.nseg += sw.nseg;
    w.nraw += ((int64_t)sw.len + 15) & ~(int64_t)15;
    w.nchunks += (int64_t)sw.chunks.size();
    const size_t base = w.bounds.back();
    for (size_t k = 1; k < bd.size(); ++k) w.bounds.push_back(base + bd[k]);
    scan_bytes += sw.len;
    coded += sw.comp < 0 ? f.num_blocks : sw.units;
    if (sw.prog >= 0) {
      const size_t lev = (size_t)w.level[(size_t)sw.prog];
      if (w.lev_count.size() <= lev) w.lev_count.resize(lev + 1, 0);
      w.lev_count[lev] += sw.nseg;
    }
  }
  w.nwords = 0;
  w.nitems = 0;
  w.S = b.tile_bits;
  w.C = b.ck_bits;
  w.items = b.item_tiles;
  const uint64_t bpb = coded ? scan_bytes * 8 / (uint64_t)coded : 0;
  for (const TuneEntry& t : b.tune)
    if (scan_bytes <= t.max_scan_bytes && (!t.max_bpb || bpb <= t.max_bpb)) {
      w.S = t.tile_bits;
      w.C = t.ck_bits;
      w.items = t.item_tiles;
      break;
    }
  if (!w.C) w.C = std::max<uint32_t>(64, w.S / 16);
  const uint32_t per_first = w.items, per_next = w.items - 1;
  const TileGeom g = tile_geom(w.S, w.C);
  for (int64_t s = 0; s < w.nseg; ++s) {
    const size_t dlen = w.bounds[s + 1] - w.bounds[s];
    w.nwords += seg_words(dlen);
    if (st.progressive) continue;  // decoded by k_prog: no entropy work items
    for_items(ntiles_of(dlen, w.S), per_first, per_next, [&](uint32_t t0, uint32_t n) {
      ++w.nitems;
      w.nent += (int64_t)item_words(n + (t0 > 0 ? 1u : 0u), g.rs, g.rd, g.nck);
    });
  }
  const int G = std::max(1, kPixBlocks / f.blocks_per_mcu);
  w.npix = (int64_t)f.mcus_y * ((f.mcus_x + G - 1) / G);
  return 0;
}

// Pass 3 (per image, in parallel): write the image's words, descriptors and
// work items at the offsets of pass 2.
void image_pass3(Plan* b, int index, ImageRec& r, ImageWork& w, const uint8_t* file) {
  const cjd_frame_info& f = r.f;
  const int nh = f.blocks_per_mcu;
  ImageDesc d{};
  d.width = f.width;
  d.height = f.height;
  d.ncomp = f.num_channels;
  d.mcus_x = f.mcus_x;
  d.mcus_y = f.mcus_y;
  d.nslots = f.blocks_per_mcu;
  d.mcu_w = f.mcu_width;
  d.mcu_h = f.mcu_height;
  d.table_base = (int32_t)(w.tset0 * 4);
  d.qtab_base = (int32_t)(w.desc * 4);
  int slot = 0;
  for (int c = 0; c < f.num_channels; ++c) {
    d.comp_first_slot[c] = slot;
    d.comp_dq[c] = (uint8_t)f.dq_id[c];
    d.comp_sx[c] = (uint8_t)f.samples_x[c];
    d.comp_sy[c] = (uint8_t)f.samples_y[c];
    int ww = f.comp_width[c], hh = f.comp_height[c], xs = 0, ys = 0;
    while (ww < f.width) { ww <<= 1; ++xs; }   // CPUonly/src/pixelTransformCPU.c:113-114
    while (hh < f.height) { hh <<= 1; ++ys; }
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
  // a Huffman table set per scan: the four tables in effect at it
  for (size_t k = 0; k < w.scans.size(); ++k) {
    const ScanWork& sw = w.scans[k];
    const size_t ts = (size_t)w.tset0 + k;
    uint16_t* lut = &b->luts[ts * kLutEntries];
    int pool = 0;
    for (int t = 0; t < 4; ++t) build_lut(sw.bits[t], sw.vals[t], t >= 2, lut, lut + t * kL1Size, pool);
    for (int e = 4 * kL1Size + pool; e < kLutEntries; ++e) lut[e] = 0;
    for (int t = 0; t < 4; ++t) build_table(sw.bits[t], sw.vals[t], b->tables[ts * 4 + t]);
  }
  // quant tables by id (one-component scans: each component's as latched at
  // its scan, under id c -- see parse_until_scan)
  for (int t = 0; t < 4; ++t) {
    uint16_t* q = &b->qtabs[(size_t)(w.desc * 4 + t) * 64];
    for (int k = 0; k < 64; ++k) q[kZZ[k]] = f.dq_tables16[t][k];
  }
  d.coef_base = w.coef0;
  d.nblocks = f.num_blocks;
  d.out_offset = w.out0;
  for (int c = 0; c < 3; ++c) d.plane_offset[c] = w.plane0[c];
  d.error_index = index;
  b->images[(size_t)w.desc] = d;

  // words: 0xFF padding, then the unstuffed data written straight in; or,
  // for the device unstuffer, the stuffed scan and its chunk states
  std::vector<size_t> dst_off((size_t)w.nseg);
  int64_t wo = w.word0;
  uint32_t* words = b->wptr;
  for (int64_t s = 0; s < w.nseg; ++s) {
    const size_t dlen = w.bounds[s + 1] - w.bounds[s];
    const int64_t nw = seg_words(dlen);
    if (words) std::fill(words + wo, words + wo + nw, 0xFFFFFFFFu);
    dst_off[(size_t)s] = (size_t)wo * 4;
    wo += nw;
  }
  for (const ScanWork& sw : w.scans) {
    if (b->unstuff_device) {
      memcpy(b->rptr + w.raw0 + sw.raw_rel, file + sw.off, (size_t)sw.len);
      for (size_t k = 0; k < sw.chunks.size(); ++k) {
        UnstuffChunk c = sw.chunks[k];
        c.src += (uint64_t)(w.raw0 + sw.raw_rel);
        c.seg += (uint32_t)(w.seg0 + sw.seg_rel);
        b->chunks[(size_t)(w.chunk0 + sw.chunk_rel) + k] = c;
      }
    } else {
      unstuff_segments(file + sw.off, (size_t)sw.len, sw.ri > 0, nullptr, reinterpret_cast<uint8_t*>(words),
                       dst_off.data() + sw.seg_rel, nullptr);
    }
  }

  // segments, tiles and entropy work items: one lane per tile; an item after
  // the first also decodes its lead-in tile
  const uint32_t per_first = w.items, per_next = w.items - 1;
  const TileGeom g = tile_geom(w.S, w.C);
  int64_t item = w.item0;
  int64_t eo = w.ent0;
  size_t k = 0;
  for (int64_t s = 0; s < w.nseg; ++s) {
    while (s >= w.scans[k].seg_rel + w.scans[k].nseg) ++k;
    const ScanWork& sw = w.scans[k];
    const size_t dlen = w.bounds[s + 1] - w.bounds[s];
    SegmentDesc sd{};
    sd.word_base = dst_off[(size_t)s] / 4;
    sd.nbits = (uint32_t)(dlen * 8);
    sd.nwords = (uint32_t)seg_words(dlen);
    sd.image = (uint32_t)w.desc;
    sd.tile_bits = w.S;
    sd.ck_bits = w.C;
    sd.nck = g.nck;
    sd.rs = g.rs;
    sd.rd = g.rd;
    sd.ntiles = ntiles_of(dlen, w.S);
    // the segment's MCUs of its scan (one block each in a one-component scan)
    const int64_t per = sw.ri > 0 ? sw.ri : sw.units;
    const int64_t u0 = (s - sw.seg_rel) * per;
    const int64_t units = std::min<int64_t>(sw.units - u0, per);
    sd.tset = (uint32_t)(w.tset0 + (int64_t)k);
    if (sw.prog >= 0) {
      const ScanInfo& sc = w.st.scans[(size_t)sw.prog];
      sd.prog = 1;
      sd.nblocks = (uint32_t)units;
      sd.first_block = u0;
      sd.ntiles = 0;
      ProgTask t{};
      t.seg = (uint32_t)(w.seg0 + s);
      t.u0 = (uint32_t)u0;
      t.nunits = (uint32_t)units;
      t.bw = (uint32_t)sw.bw;
      t.ncomp = (uint8_t)sc.ncomp;
      t.ss = (uint8_t)sc.ss;
      t.se = (uint8_t)sc.se;
      t.ah = (uint8_t)sc.ah;
      t.al = (uint8_t)sc.al;
      int bpu = 0;
      for (int i = 0; i < sc.ncomp; ++i) {
        const int c = sc.comps[i];
        t.comps[i] = (uint8_t)c;
        t.dcs[i] = (uint8_t)sc.dcs[i];
        const int n = sc.ncomp > 1 ? f.samples_x[c] * f.samples_y[c] : 1;
        for (int q = 0; q < n && sc.ncomp > 1; ++q) {
          t.slots |= (uint64_t)(d.comp_first_slot[c] + q) << (4 * (bpu + q));
          t.cis |= (uint32_t)i << (2 * (bpu + q));
        }
        bpu += n;
      }
      t.bpu = (uint8_t)bpu;
      t.h = (uint8_t)sw.h;
      t.v = (uint8_t)sw.v;
      t.ac = (uint8_t)sc.ac_id;
      t.last = s + 1 == sw.seg_rel + sw.nseg;
      const size_t lev = (size_t)w.level[(size_t)sw.prog];
      b->ptasks[(size_t)w.lev_off[lev]++] = t;
      b->segs[(size_t)(w.seg0 + s)] = sd;
      continue;
    }
    if (sw.comp < 0) {
      sd.nblocks = (uint32_t)(units * nh);
      sd.first_block = d.coef_base + u0 * nh;
    } else {
      sd.ni = 1;
      sd.comp = (uint8_t)sw.comp;
      sd.h = (uint8_t)sw.h;
      sd.v = (uint8_t)sw.v;
      sd.dc_id = (uint8_t)sw.dc_id;
      sd.ac_id = (uint8_t)sw.ac_id;
      sd.slot0 = (uint8_t)sw.slot0;
      sd.bw = (uint32_t)sw.bw;
      sd.nblocks = (uint32_t)units;
      sd.first_block = u0;
    }
    sd.first_item = (uint32_t)item;
    uint32_t prev = ~0u;
    for_items(sd.ntiles, per_first, per_next, [&](uint32_t t0, uint32_t n) {
      WorkItem wi{};
      wi.segment = (uint32_t)(w.seg0 + s);
      wi.t0 = t0;
      wi.ntiles = n;
      wi.prev_item = prev;
      wi.ebase = (uint64_t)eo;
      eo += (int64_t)item_words(n + (t0 > 0 ? 1u : 0u), g.rs, g.rd, g.nck);
      prev = (uint32_t)item;
      b->items[(size_t)item++] = wi;
    });
    b->segs[(size_t)(w.seg0 + s)] = sd;
  }

  // pixel work items: runs of MCUs within one MCU row
  const int G = std::max(1, kPixBlocks / nh);
  int64_t pi_i = w.pix0;
  for (int my = 0; my < f.mcus_y; ++my)
    for (int mx = 0; mx < f.mcus_x; mx += G) {
      PixItem pi{};
      pi.image = (uint32_t)w.desc;
      pi.mcu_first = (uint32_t)(my * f.mcus_x + mx);
      pi.count = (uint32_t)std::min(G, f.mcus_x - mx);
      b->pix[(size_t)pi_i++] = pi;
    }
}

// A small persistent pool of host threads for the per-image passes.  run()
// calls are serialised; the caller takes part and returns once every index
// is done and no worker is still inside the job.
class Pool {
 public:
  static Pool& get() {
    static Pool p;
    return p;
  }
  void run(int n, const std::function<void(int)>& fn) {
    if (workers_.empty() || n < 2) {
      for (int i = 0; i < n; ++i) fn(i);
      return;
    }
    std::lock_guard<std::mutex> one(run_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      n_ = n;
      next_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) fn(i);
    std::unique_lock<std::mutex> lk(mu_);
    job_ = nullptr;  // no worker enters from here on
    done_.wait(lk, [&] { return running_ == 0; });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  Pool() {
    const unsigned hc = std::thread::hardware_concurrency();
    const int nw = (int)std::min<unsigned>(hc > 1 ? hc - 1 : 0, 15u);
    for (int k = 0; k < nw; ++k) workers_.emplace_back([this] { loop(); });
  }
  void loop() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || (job_ && gen_ != seen); });
      if (stop_) return;
      seen = gen_;
      const std::function<void(int)>* fn = job_;
      const int n = n_;
      ++running_;
      lk.unlock();
      for (int i; (i = next_.fetch_add(1)) < n;) (*fn)(i);
      lk.lock();
      if (--running_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, running_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace

void plan_batch(Plan& b, int n, const uint8_t* const* files, const size_t* sizes, int* errors) {
  b.recs.assign((size_t)n, ImageRec{});
  std::vector<ImageWork> work((size_t)n);
  // pass 1: parse and count, per image in parallel
  Pool::get().run(n, [&](int i) {
    ImageRec& r = b.recs[(size_t)i];
    r.error = image_pass1(b, r, work[(size_t)i], files[i], sizes[i]);
  });
  // pass 2: offsets, in image order
  int64_t nimg = 0, nseg = 0, nitems = 0, npix = 0, nwords = 0, nblocks = 0, nent = 0, nraw = 0, nchunks = 0;
  int64_t ntset = 0;
  b.total_pixels = b.scan_bytes = 0;
  b.out_bytes = 0;
  for (int i = 0; i < n; ++i) {
    ImageRec& r = b.recs[(size_t)i];
    if (errors) errors[i] = r.error;
    if (r.error) continue;
    ImageWork& w = work[(size_t)i];
    const cjd_frame_info& f = r.f;
    w.desc = nimg++;
    w.seg0 = nseg;
    w.item0 = nitems;
    w.pix0 = npix;
    w.word0 = nwords;
    w.coef0 = nblocks;
    w.ent0 = nent;
    nent += w.nent;
    w.raw0 = nraw;
    nraw += w.nraw;
    w.chunk0 = nchunks;
    nchunks += w.nchunks;
    w.tset0 = ntset;
    ntset += (int64_t)w.scans.size();
    nseg += w.nseg;
    nitems += w.nitems;
    npix += w.npix;
    nwords += w.nwords;
    nblocks += f.num_blocks;
    const uint64_t ob = (uint64_t)f.width * f.height * (f.num_channels == 3 ? 3 : 1);
    w.out0 = r.out_offset = b.out_bytes;
    r.out_bytes = ob;
    b.out_bytes += (ob + 255) & ~255ull;
    if (b.want_planes)
      for (int c = 0; c < f.num_channels; ++c) {
        const uint64_t pb = (uint64_t)f.comp_stride[c] * f.comp_plane_height[c];
        w.plane0[c] = r.plane_offset[c] = b.out_bytes;
        b.out_bytes += (pb + 255) & ~255ull;
      }
    r.coef_base = w.coef0;
    r.desc = (uint32_t)w.desc;
    r.seg_first = (uint32_t)w.seg0;
    r.seg_count = (uint32_t)w.nseg;
    r.item_first = (uint32_t)w.item0;
    r.item_count = (uint32_t)w.nitems;
    r.pix_first = (uint32_t)w.pix0;
    r.pix_count = (uint32_t)w.npix;
    r.pixels = (int64_t)f.width * f.height;
    b.total_pixels += r.pixels;
    b.scan_bytes += (int64_t)w.bounds.back();
  }
  // progressive tasks, level by level across the batch; each image's tasks
  // of a level follow those of the images before it
  size_t nlev = 0;
  for (int i = 0; i < n; ++i)
    if (!b.recs[(size_t)i].error) nlev = std::max(nlev, work[(size_t)i].lev_count.size());
  b.plevel.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (b.recs[(size_t)i].error) continue;
    const ImageWork& w = work[(size_t)i];
    for (size_t l = 0; l < w.lev_count.size(); ++l) b.plevel[l + 1] += (uint32_t)w.lev_count[l];
  }
  for (size_t l = 0; l < nlev; ++l) b.plevel[l + 1] += b.plevel[l];
  std::vector<int64_t> cursor(b.plevel.begin(), b.plevel.end());
  b.pimgs.clear();
  b.prog_blocks = 0;
  for (int i = 0; i < n; ++i) {
    const ImageRec& r = b.recs[(size_t)i];
    if (r.error || !r.f.progressive) continue;
    ImageWork& w = work[(size_t)i];
    w.lev_off.assign(w.lev_count.size(), 0);
    for (size_t l = 0; l < w.lev_count.size(); ++l) {
      w.lev_off[l] = cursor[l];
      cursor[l] += w.lev_count[l];
    }
    b.pimgs.push_back(ProgImage{r.coef_base, r.f.num_blocks, b.prog_blocks});
    b.prog_blocks += r.f.num_blocks;
  }
  b.ptasks.resize(b.plevel.back());
  if (nlev == 0) b.plevel.clear();
  b.total_blocks = nblocks;
  b.entry_words = nent;
  b.images.resize((size_t)nimg);
  b.segs.resize((size_t)nseg);
  b.items.resize((size_t)nitems);
  b.pix.resize((size_t)npix);
  b.tables.resize((size_t)ntset * 4);
  b.luts.resize((size_t)ntset * kLutEntries);
  b.qtabs.resize((size_t)nimg * 4 * 64);
  b.nwords = (size_t)nwords;
  b.chunks.resize((size_t)nchunks);
  b.nraw = b.unstuff_device ? (size_t)nraw : 0;
  b.wptr = nullptr;
  b.rptr = nullptr;
  if (b.unstuff_device) {
    b.words.clear();
    b.rptr = b.raw_store ? b.raw_store(b.nraw) : nullptr;
    if (!b.rptr) {
      b.raw.resize(b.nraw);
      b.rptr = b.raw.data();
    }
  } else {
    b.wptr = b.word_store ? b.word_store(b.nwords) : nullptr;
    if (b.wptr) {
      b.words.clear();
    } else {
      b.words.resize(b.nwords);
      b.wptr = b.words.data();
    }
  }
  // pass 3: fill, per image in parallel
  Pool::get().run(n, [&](int i) {
    ImageRec& r = b.recs[(size_t)i];
    if (!r.error) image_pass3(&b, i, r, work[(size_t)i], files[i]);
  });
}

}  // namespace cjd
