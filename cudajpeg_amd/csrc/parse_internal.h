// Host parser state shared by the C ABI (cjd_parse, batch staging) and the
// reference-style JPGReader API.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/cudajpeg_amd.h"

namespace cjd {

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
};

void parse_sof(ParseState& st);
void parse_dht(ParseState& st);
void parse_dri(ParseState& st);
void parse_dqt(ParseState& st);
void parse_sos(ParseState& st);
int parse_until_scan(ParseState& st, const uint8_t* data, size_t size, cjd_frame_info* f, int ext = 0);

}  // namespace cjd
