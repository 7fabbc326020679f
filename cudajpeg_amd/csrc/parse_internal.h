// Host parser state shared by the C ABI (cjd_parse, batch staging) and the
// reference-style JPGReader API.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/cudajpeg_amd.h"

namespace cjd {

// A one-component scan (CJD_EXT_NONINTERLEAVED) and the state in effect at
// its SOS.
struct ScanInfo {
  int comp = 0, dc_id = 0, ac_id = 2, restart_interval = 0;
  uint64_t offset = 0, length = 0;  // its stuffed bytes: file[offset, offset + length)
  uint8_t huff_bits[4][16];
  uint8_t huff_vals[4][256];
  uint16_t dq[64];                  // the component's quant table when the scan began (zigzag)
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
  int nscans = 0;  // one-component scans parsed (CJD_EXT_NONINTERLEAVED)
  ScanInfo scans[3];
};

void parse_sof(ParseState& st);
void parse_dht(ParseState& st);
void parse_dri(ParseState& st);
void parse_dqt(ParseState& st);
void parse_sos(ParseState& st);
// CJD_EXT_NONINTERLEAVED: the SOS of a one-component scan; leaves st.pos at
// the marker after the scan's data.
void parse_sos_one(ParseState& st);
int parse_until_scan(ParseState& st, const uint8_t* data, size_t size, cjd_frame_info* f, int ext = 0);

}  // namespace cjd
