// Host marker / table parser of the MI355X baseline-JPEG decoder.
//
// Same checks, same order and same error codes as the reference parser
// (src/format.cu:113-442 and the SOS header of src/decodeScanGPU.cu:124-141),
// operating on an in-memory file.  The 65536-entry Huffman LUTs of the
// reference (src/format.cu:371-391) are not built here: DHT data is kept as
// BITS/HUFFVAL and turned into compact device tables by the engine.
#include <string.h>

#include "../../include/cudajpeg_amd.h"
#include "parse_internal.h"

namespace cjd {

static inline unsigned read16(const uint8_t* p) { return (unsigned)((p[0] << 8) | p[1]); }

// most scans of a progressive file (T.81 sets no limit; libjpeg's scripts use
// 6 to 10)
constexpr int kMaxScans = 64;

#define PFAIL(e)              \
  do {                        \
    st.error = (e);           \
    st.error_line = __LINE__; \
    return;                   \
  } while (0)

// src/format.cu:242-354
void parse_sof(ParseState& st) {
  cjd_frame_info& f = *st.f;
  const uint8_t* block = st.pos;
  const unsigned block_len = read16(block);
  if (block_len < 9 || block + block_len >= st.end) PFAIL(CJD_SYNTAX_ERROR);
  if (block[2] != 8) PFAIL(CJD_UNSUPPORTED_ERROR);
  f.height = (int)read16(block + 3);
  f.width = (int)read16(block + 5);
  if (!f.width || !f.height) PFAIL(CJD_SYNTAX_ERROR);
  f.num_channels = block[7];
  if (f.num_channels != 1 && f.num_channels != 3) PFAIL(CJD_UNSUPPORTED_ERROR);
  if (block_len < 8u + (unsigned)(f.num_channels * 3)) PFAIL(CJD_SYNTAX_ERROR);
  block += 8;
  int sxm = 0, sym = 0;
  for (int i = 0; i < f.num_channels; ++i, block += 3) {
    f.comp_id[i] = block[0];
    f.samples_x[i] = block[1] >> 4;
    f.samples_y[i] = block[1] & 0xF;
    f.dq_id[i] = block[2];
    if (!f.samples_x[i] || !f.samples_y[i] || f.dq_id[i] > 3) PFAIL(CJD_SYNTAX_ERROR);
    if ((f.samples_x[i] & (f.samples_x[i] - 1)) || (f.samples_y[i] & (f.samples_y[i] - 1)))
      PFAIL(CJD_UNSUPPORTED_ERROR);
    if (f.samples_x[i] > sxm) sxm = f.samples_x[i];
    if (f.samples_y[i] > sym) sym = f.samples_y[i];
  }
  if (f.num_channels == 1) {  // src/format.cu:281-284
    f.samples_x[0] = sxm = 1;
    f.samples_y[0] = sym = 1;
  }
  f.mcu_width = sxm << 3;
  f.mcu_height = sym << 3;
  f.mcus_x = (f.width + f.mcu_width - 1) / f.mcu_width;
  f.mcus_y = (f.height + f.mcu_height - 1) / f.mcu_height;
  f.blocks_per_mcu = 0;
  for (int i = 0; i < f.num_channels; ++i) {
    f.comp_width[i] = (f.width * f.samples_x[i] + sxm - 1) / sxm;
    f.comp_height[i] = (f.height * f.samples_y[i] + sym - 1) / sym;
    f.comp_stride[i] = f.mcus_x * (f.samples_x[i] << 3);
    f.comp_plane_height[i] = f.mcus_y * (f.samples_y[i] << 3);
    if ((f.comp_width[i] < 3 && f.samples_x[i] != sxm) || (f.comp_height[i] < 3 && f.samples_y[i] != sym))
      PFAIL(CJD_UNSUPPORTED_ERROR);
    f.blocks_per_mcu += f.samples_x[i] * f.samples_y[i];
  }
  if (f.num_channels == 3) {  // src/format.cu:346-350
    if (f.comp_width[1] != f.comp_width[2] || f.comp_height[1] != f.comp_height[2] ||
        f.comp_width[0] != f.width || f.comp_height[0] != f.height)
      PFAIL(CJD_UNSUPPORTED_ERROR);
  }
  f.num_blocks = (int64_t)f.mcus_x * f.mcus_y * f.blocks_per_mcu;
  st.have_sof = true;
  st.pos += block_len;
}

// src/format.cu:357-400 (table id = Tc*2 + Th)
void parse_dht(ParseState& st) {
  cjd_frame_info& f = *st.f;
  const uint8_t* pos = st.pos;
  const unsigned block_len = read16(pos);
  const uint8_t* block_end = pos + block_len;
  if (block_end > st.end) PFAIL(CJD_SYNTAX_ERROR);
  pos += 2;
  while (pos < block_end) {
    const uint8_t val = pos[0];
    if (val & 0xEC) PFAIL(CJD_SYNTAX_ERROR);
    if (val & 0x02) PFAIL(CJD_UNSUPPORTED_ERROR);
    const int id = (val | (val >> 3)) & 3;
    const uint8_t* tuple = pos + 17;
    int remain = 65536, n = 0;
    uint8_t bits[16];
    uint8_t vals[256];
    for (int l = 1; l <= 16; ++l) {
      const int count = pos[l];
      bits[l - 1] = (uint8_t)count;
      if (!count) continue;
      if (tuple + count > block_end) PFAIL(CJD_SYNTAX_ERROR);
      remain -= count << (16 - l);
      if (remain < 0) PFAIL(CJD_SYNTAX_ERROR);
      // HUFFVAL holds at most 256 symbols (T.81 B.2.4.2); the Kraft sum
      // alone admits up to 2040 codes of lengths 9..16
      if (n + count > 256) PFAIL(CJD_SYNTAX_ERROR);
      for (int i = 0; i < count; ++i) vals[n++] = *tuple++;
    }
    memcpy(f.huff_bits[id], bits, 16);
    memset(f.huff_vals[id], 0, 256);
    memcpy(f.huff_vals[id], vals, (size_t)n);
    pos = tuple;
  }
  if (pos != block_end) PFAIL(CJD_SYNTAX_ERROR);
  st.pos = block_end;
}

// src/format.cu:404-421
void parse_dri(ParseState& st) {
  const unsigned block_len = read16(st.pos);
  const uint8_t* block_end = st.pos + block_len;
  if (block_len < 2 || block_end > st.end) PFAIL(CJD_SYNTAX_ERROR);
  st.f->restart_interval = (int)read16(st.pos + 2);
  st.pos = block_end;
}

// src/format.cu:424-442
// src/format.cu:424-442; with CJD_EXT_DQT16 a table may also have 16-bit
// entries (Pq = 1, T.81 B.2.4.1: 128 bytes, big-endian).
void parse_dqt(ParseState& st) {
  const unsigned block_len = read16(st.pos);
  const uint8_t* block_end = st.pos + block_len;
  if (block_end > st.end) PFAIL(CJD_SYNTAX_ERROR);
  const uint8_t* pos = st.pos + 2;
  const bool wide_ok = (st.ext & CJD_EXT_DQT16) != 0;
  while (pos + 65 <= block_end) {
    const uint8_t id = pos[0];
    const bool wide = wide_ok && (id & 0xF0) == 0x10;
    if ((id & 0xFC) && !wide) PFAIL(CJD_SYNTAX_ERROR);
    if ((id & 0x0C) || (wide && pos + 129 > block_end)) PFAIL(CJD_SYNTAX_ERROR);
    const int t = id & 3;
    for (int k = 0; k < 64; ++k) {
      const uint16_t v = wide ? (uint16_t)((pos[1 + 2 * k] << 8) | pos[2 + 2 * k]) : pos[1 + k];
      st.f->dq_tables16[t][k] = v;
      st.f->dq_tables[t][k] = (uint8_t)(v > 255 ? 255 : v);
    }
    pos += wide ? 129 : 65;
  }
  if (pos != block_end) PFAIL(CJD_SYNTAX_ERROR);
  st.pos = block_end;
}

// src/decodeScanGPU.cu:124-141
void parse_sos(ParseState& st) {
  cjd_frame_info& f = *st.f;
  const uint8_t* pos = st.pos;
  const unsigned header_len = read16(pos);
  if (pos + header_len > st.end) PFAIL(CJD_SYNTAX_ERROR);
  pos += 2;
  if (header_len < 4u + 2u * (unsigned)f.num_channels) PFAIL(CJD_SYNTAX_ERROR);
  if (*(pos++) != f.num_channels) PFAIL(CJD_UNSUPPORTED_ERROR);
  for (int i = 0; i < f.num_channels; ++i, pos += 2) {
    if (pos[0] != f.comp_id[i]) PFAIL(CJD_SYNTAX_ERROR);
    if (pos[1] & 0xEE) PFAIL(CJD_SYNTAX_ERROR);
    f.dc_id[i] = pos[1] >> 4;
    f.ac_id[i] = (pos[1] & 1) | 2;
  }
  if (pos[0] || pos[1] != 63 || pos[2]) PFAIL(CJD_UNSUPPORTED_ERROR);
  st.pos += header_len;
  // The scan runs to the EOI: the reference unstuffs everything up to
  // jpg->end (src/decodeScanGPU.cu:13).
  f.scan_offset = (uint64_t)(st.pos - st.buf);
  f.scan_length = (uint64_t)(st.end - st.pos);
}

// The data of a scan that starts at d: to the first 0xFF followed by neither
// 0x00 nor an RSTn (T.81 B.1.1.2, B.1.1.5), at the latest to the EOI; st.pos
// goes there.  The first scan's bounds are the frame's scan_offset/length.
static void scan_data(ParseState& st, ScanInfo& sc, const uint8_t* d) {
  const uint8_t* e = d;
  for (;;) {
    const void* ff = e < st.end ? memchr(e, 0xFF, (size_t)(st.end - e)) : nullptr;
    if (!ff) {
      e = st.end;
      break;
    }
    e = static_cast<const uint8_t*>(ff);  // e < st.end: e[1] is in the file
    if (e[1] != 0x00 && (e[1] & 0xF8) != 0xD0) break;
    e += 2;
  }
  sc.offset = (uint64_t)(d - st.buf);
  sc.length = (uint64_t)(e - d);
  if (st.nscans == 1) {
    st.f->scan_offset = sc.offset;
    st.f->scan_length = sc.length;
  }
  st.pos = e;
}

// T.81 B.2.3 with Ns = 1 (src/decodeScanGPU.cu:124-131 rejects it): the
// scan's component, its tables and the restart interval in effect, the
// component's quant table as it is now; the scan's data runs to the first
// 0xFF followed by neither 0x00 nor an RSTn (T.81 B.1.1.2, B.1.1.5), at the
// latest to the EOI.
void parse_sos_one(ParseState& st) {
  cjd_frame_info& f = *st.f;
  const uint8_t* pos = st.pos;
  const unsigned header_len = read16(pos);
  if (header_len != 8 || pos + header_len > st.end) PFAIL(CJD_SYNTAX_ERROR);
  pos += 3;
  int c = 0;
  while (c < f.num_channels && f.comp_id[c] != pos[0]) ++c;
  if (c == f.num_channels || st.nscans >= f.num_channels) PFAIL(CJD_SYNTAX_ERROR);
  for (int k = 0; k < st.nscans; ++k)
    if (st.scans[k].comp == c) PFAIL(CJD_SYNTAX_ERROR);  // each component in one scan
  if (pos[1] & 0xEE) PFAIL(CJD_SYNTAX_ERROR);
  if (pos[2] || pos[3] != 63 || pos[4]) PFAIL(CJD_UNSUPPORTED_ERROR);
  st.scans.emplace_back();
  ScanInfo& sc = st.scans[(size_t)st.nscans++];
  sc.comp = c;
  sc.dc_id = pos[1] >> 4;
  sc.ac_id = (pos[1] & 1) | 2;
  sc.restart_interval = f.restart_interval;
  memcpy(sc.huff_bits, f.huff_bits, sizeof(sc.huff_bits));
  memcpy(sc.huff_vals, f.huff_vals, sizeof(sc.huff_vals));
  memcpy(sc.dq, f.dq_tables16[f.dq_id[c]], sizeof(sc.dq));
  scan_data(st, sc, st.pos + header_len);
}

// T.81 B.2.3 in a progressive frame (G.1.1.1), which the reference never
// reaches (src/format.cu:198-200 rejects SOF2): Ns components in frame
// order; a DC scan (Ss = 0) has Se = 0, an AC scan 1 <= Ss <= Se <= 63 and
// one component; Al <= 13 and a refinement (Ah > 0) has Al = Ah - 1.  The
// band's previous scan must have ended at Ah (cbits), a first scan must be
// the band's first, and an AC scan needs the component's DC begun.  Each
// component's quant table is latched at its first scan.
void parse_sos_prog(ParseState& st) {
  cjd_frame_info& f = *st.f;
  if (!st.have_sof) PFAIL(CJD_SYNTAX_ERROR);
  const uint8_t* pos = st.pos;
  const unsigned header_len = read16(pos);
  if (header_len < 3 || pos + header_len > st.end) PFAIL(CJD_SYNTAX_ERROR);
  const int ns = pos[2];
  if (ns < 1 || ns > f.num_channels || header_len != 6u + 2u * (unsigned)ns) PFAIL(CJD_SYNTAX_ERROR);
  if (st.nscans >= kMaxScans) PFAIL(CJD_UNSUPPORTED_ERROR);
  ScanInfo sc;
  sc.ncomp = ns;
  pos += 3;
  int last = -1;
  for (int i = 0; i < ns; ++i, pos += 2) {
    int c = last + 1;
    while (c < f.num_channels && f.comp_id[c] != pos[0]) ++c;
    if (c == f.num_channels) PFAIL(CJD_SYNTAX_ERROR);  // unknown, repeated or out of frame order
    if (pos[1] & 0xEE) PFAIL(CJD_SYNTAX_ERROR);
    sc.comps[i] = c;
    sc.dcs[i] = pos[1] >> 4;
    sc.acs[i] = (pos[1] & 1) | 2;
    last = c;
  }
  sc.ss = pos[0];
  sc.se = pos[1];
  sc.ah = pos[2] >> 4;
  sc.al = pos[2] & 15;
  if (sc.ss == 0 ? sc.se != 0 : (sc.se < sc.ss || sc.se > 63 || ns != 1)) PFAIL(CJD_SYNTAX_ERROR);
  if (sc.al > 13 || (sc.ah && sc.al != sc.ah - 1)) PFAIL(CJD_SYNTAX_ERROR);
  for (int i = 0; i < ns; ++i) {
    const int8_t* cb = st.cbits[sc.comps[i]];
    if (sc.ss > 0 && cb[0] < 0) PFAIL(CJD_SYNTAX_ERROR);
    for (int k = sc.ss; k <= sc.se; ++k)
      if (cb[k] != (sc.ah ? sc.ah : -1)) PFAIL(CJD_SYNTAX_ERROR);
  }
  for (int i = 0; i < ns; ++i) {
    const int c = sc.comps[i];
    for (int k = sc.ss; k <= sc.se; ++k) st.cbits[c][k] = (int8_t)sc.al;
    if (!st.latched[c]) {
      st.latched[c] = true;
      memcpy(st.comp_dq[c], f.dq_tables16[f.dq_id[c]], sizeof(st.comp_dq[c]));
    }
  }
  sc.comp = sc.comps[0];
  sc.dc_id = sc.dcs[0];
  sc.ac_id = sc.acs[0];
  sc.restart_interval = f.restart_interval;
  memcpy(sc.huff_bits, f.huff_bits, sizeof(sc.huff_bits));
  memcpy(sc.huff_vals, f.huff_vals, sizeof(sc.huff_vals));
  st.scans.push_back(sc);
  ++st.nscans;
  scan_data(st, st.scans.back(), st.pos + header_len);
}

// src/format.cu:113-215 without the file read; stops after the SOS header
// (or at the EOI) and leaves st.pos there.  With CJD_EXT_NONINTERLEAVED, a
// frame of several components whose first scan has one component is parsed
// to the EOI, scan by scan.
int parse_until_scan(ParseState& st, const uint8_t* data, size_t size, cjd_frame_info* f, int ext) {
  memset(f, 0, sizeof(*f));
  st.f = f;
  st.ext = ext;
  st.buf = data;
  st.error = 0;
  st.error_line = 0;
  st.have_sof = false;
  st.at_scan = false;
  st.nscans = 0;
  st.scans.clear();
  st.progressive = false;
  memset(st.cbits, -1, sizeof(st.cbits));
  for (bool& l : st.latched) l = false;
  if (size < 4 || data[0] != 0xFF || data[1] != 0xD8 || data[size - 2] != 0xFF || data[size - 1] != 0xD9) {
    st.error = CJD_FILE_ERROR;
    st.error_line = __LINE__;
    return st.error;
  }
  if (size < 6) {
    st.error = CJD_SYNTAX_ERROR;
    st.error_line = __LINE__;
    return st.error;
  }
  st.end = data + size - 2;
  st.pos = data + 2;
  while (!st.error) {
    if (st.pos > st.end || st.pos[0] != 0xFF) {
      st.error = CJD_SYNTAX_ERROR;
      st.error_line = __LINE__;
      break;
    }
    st.pos += 2;
    const uint8_t m = st.pos[-1];
    switch (m) {
      case 0xC0: parse_sof(st); break;
      case 0xC2:  // src/format.cu:198-200 rejects it as an unknown marker
        if (!(st.ext & CJD_EXT_PROGRESSIVE)) {
          st.error = CJD_SYNTAX_ERROR;
          st.error_line = __LINE__;
          break;
        }
        parse_sof(st);
        st.progressive = true;
        break;
      case 0xC4: parse_dht(st); break;
      case 0xDB: parse_dqt(st); break;
      case 0xDD: parse_dri(st); break;
      case 0xDA:
        if (st.progressive) {
          parse_sos_prog(st);
          break;
        }
        if ((st.ext & CJD_EXT_NONINTERLEAVED) && f->num_channels > 1 && st.pos + 3 <= st.end && st.pos[2] == 1) {
          parse_sos_one(st);
          break;
        }
        if (st.nscans) {  // an interleaved scan after one-component scans
          st.error = CJD_SYNTAX_ERROR;
          st.error_line = __LINE__;
          break;
        }
        parse_sos(st);
        if (!st.error) st.at_scan = true;
        break;
      case 0xFE: st.pos += read16(st.pos); break;
      case 0xD9: break;
      default:
        if ((m & 0xF0) == 0xE0) st.pos += read16(st.pos);
        else {
          st.error = CJD_SYNTAX_ERROR;
          st.error_line = __LINE__;
        }
    }
    if (st.error || st.at_scan || m == 0xD9) break;
  }
  if (!st.error && st.progressive) {
    // each component's quant table as latched at its first scan, listed
    // under table id c (a component no scan codes keeps zero coefficients)
    st.at_scan = st.nscans > 0;
    for (int c = 0; c < f->num_channels; ++c) {
      const uint16_t* q = st.latched[c] ? st.comp_dq[c] : f->dq_tables16[f->dq_id[c]];
      uint16_t t[64];
      memcpy(t, q, sizeof(t));
      f->dq_id[c] = c;
      for (int i = 0; i < 64; ++i) {
        f->dq_tables16[c][i] = t[i];
        f->dq_tables[c][i] = (uint8_t)(t[i] > 255 ? 255 : t[i]);
      }
    }
    f->progressive = 1;
  } else if (!st.error && st.nscans) {
    if (st.nscans != f->num_channels) {  // a component no scan codes
      st.error = CJD_SYNTAX_ERROR;
      st.error_line = __LINE__;
    } else {
      // each component's quant table as latched at its scan, listed under
      // table id c (a DQT between scans may have redefined the id it named)
      st.at_scan = true;
      for (int k = 0; k < st.nscans; ++k) {
        const int c = st.scans[k].comp;
        f->dq_id[c] = c;
        for (int i = 0; i < 64; ++i) {
          const uint16_t v = st.scans[k].dq[i];
          f->dq_tables16[c][i] = v;
          f->dq_tables[c][i] = (uint8_t)(v > 255 ? 255 : v);
        }
      }
    }
  }
  f->num_scans = st.nscans ? st.nscans : 1;
  f->error_line = st.error_line;
  return st.error;
}

}  // namespace cjd

extern "C" int cjd_parse_ext(const uint8_t* data, size_t size, int ext, cjd_frame_info* out) {
  cjd::ParseState st;
  int e = cjd::parse_until_scan(st, data, size, out, ext);
  if (!e && !st.at_scan) e = CJD_SYNTAX_ERROR;  // EOI before any scan
  if (!e && !st.have_sof) e = CJD_SYNTAX_ERROR;
  out->error_line = st.error_line;
  return e;
}

extern "C" int cjd_parse(const uint8_t* data, size_t size, cjd_frame_info* out) {
  return cjd_parse_ext(data, size, 0, out);
}

extern "C" const char* cjd_error_name(int code) {
  switch (code) {  // src/format.cu:62-78
    case CJD_NO_ERROR: return "NO_ERROR";
    case CJD_SYNTAX_ERROR: return "SYNTAX_ERROR";
    case CJD_UNSUPPORTED_ERROR: return "UNSUPPORTED_ERROR";
    case CJD_OOM_ERROR: return "OOM_ERROR";
    case CJD_DEVICE_MEM_ERROR: return "DEVICE_MEM_ERROR";
    case CJD_FILE_ERROR: return "FILE_ERROR";
    case CJD_KERNEL_LAUNCH_ERROR: return "KERNEL_LAUNCH_ERROR";
    case CJD_PROGRAMMER_ERROR: return "PROGRAMMER_ERROR";
    default: return "UNRECOGNISED_ERROR";
  }
}
