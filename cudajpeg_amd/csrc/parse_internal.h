// Host parser state shared by the C ABI (cjd_parse, batch staging) and the
// reference-style JPGReader API.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/cudajpeg_amd.h"

namespace cjd {

// A scan of a multi-scan file -- a one-component scan (CJD_EXT_NONINTERLEAVED)
// or a progressive scan (CJD_EXT_PROGRESSIVE) -- and the state in effect at
// its SOS.
struct ScanInfo {
  int comp = 0, dc_id = 0, ac_id = 2, restart_interval = 0;  // comp: its (first) component
  uint64_t offset = 0, length = 0;  // its stuffed bytes: file[offset, offset + length)
  uint8_t huff_bits[4][16];
  uint8_t huff_vals[4][256];
  uint16_t dq[64];                  // the component's quant table when the scan began (zigzag)
  // progressive scans: the components (frame indexes, frame order) with
  // their DC / AC table ids, the spectral band Ss..Se and the successive
  // approximation bits Ah, Al
  int ncomp = 1, comps[3] = {0, 0, 0}, dcs[3] = {0, 0, 0}, acs[3] = {2, 2, 2};
  int ss = 0, se = 63, ah = 0, al = 0;
};

struct ParseState {
  const uint8_t* buf = nullptr;  // start of the file
  const uint8_t* pos = nullptr;  // current marker position
  const uint8_t* end = nullptr;  // file end minus the EOI (src/format.cu:166)
  cjd_frame_info* f = nullptr;
  int error = 0;
  int error_line = 0;
  bool have_sof = false;
  bool at_scan = false;
  int ext = 0;  // CJD_EXT_* accepted
  int nscans = 0;  // scans parsed into `scans` (one-component or progressive)
  std::vector<ScanInfo> scans;
  bool progressive = false;  // SOF2 frame (CJD_EXT_PROGRESSIVE)
  // progressive: the Al of the last scan that coded coefficient k (zigzag)
  // of component c, -1 before any
  int8_t cbits[3][64];
  // progressive: each component's quant table as latched at its first scan
  bool latched[3] = {false, false, false};
  uint16_t comp_dq[3][64];
};

void parse_sof(ParseState& st);
void parse_dht(ParseState& st);
void parse_dri(ParseState& st);
void parse_dqt(ParseState& st);
void parse_sos(ParseState& st);
// CJD_EXT_NONINTERLEAVED: the SOS of a one-component scan; leaves st.pos at
// the marker after the scan's data.
void parse_sos_one(ParseState& st);
// CJD_EXT_PROGRESSIVE: the SOS of a progressive scan; likewise.
void parse_sos_prog(ParseState& st);
int parse_until_scan(ParseState& st, const uint8_t* data, size_t size, cjd_frame_info* f, int ext = 0);

}  // namespace cjd
