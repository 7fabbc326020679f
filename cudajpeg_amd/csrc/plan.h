// Host planning of a decode batch: parse, unstuff, compact Huffman tables,
// quant tables, image / segment descriptors and the entropy and pixel work
// lists.  Pure host code (no HIP calls), shared by the engine and the CPU
// emulator of the kernels.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "../../include/cudajpeg_amd.h"
#include "cjd_internal.h"
#include "unstuff_core.h"

namespace cjd {

struct ImageRec {
  int error = 0;         // parse / staging result (reference error code)
  int decode_error = 0;  // result of the last decode (reset by each check)
  int code() const { return error ? error : decode_error; }
  cjd_frame_info f{};
  int64_t pixels = 0;
  uint64_t out_offset = 0, out_bytes = 0;
  uint64_t plane_offset[3] = {~0ull, ~0ull, ~0ull};
  int64_t coef_base = 0;
  uint32_t desc = 0;  // ImageDesc index (valid when error == 0)
  uint32_t seg_first = 0, seg_count = 0;
  uint32_t item_first = 0, item_count = 0;  // entropy work items (contiguous)
  uint32_t pix_first = 0, pix_count = 0;
};

// Entropy-decoder parameters for images whose scan is at most max_scan_bytes
// and at most max_bpb scan bits per block (0 = any); the first matching entry
// of Plan::tune applies (entries sorted by size, then density).
struct TuneEntry {
  uint64_t max_scan_bytes;
  uint32_t tile_bits, ck_bits, item_tiles;
  uint32_t max_bpb = 0;
};

struct Plan {
  uint32_t tile_bits = 4096;
  uint32_t ck_bits = 0;           // checkpoint spacing C in bits, 0 = tile_bits / 16 (>= 64)
  uint32_t item_tiles = kWG;      // tiles per entropy work item (2 .. kWG)
  std::vector<TuneEntry> tune;    // per scan-size bucket overrides of the three above
  int want_planes = 0;
  int ext = 0;                    // CJD_EXT_* the parser accepts
  std::vector<ImageRec> recs;
  std::vector<ImageDesc> images;
  std::vector<SegmentDesc> segs;
  std::vector<WorkItem> items;
  std::vector<PixItem> pix;
  // Progressive images (CJD_EXT_PROGRESSIVE): one task per restart interval
  // of a scan, level by level (level l: tasks [plevel[l], plevel[l + 1])),
  // and the images' coefficient ranges.
  std::vector<ProgTask> ptasks;
  std::vector<uint32_t> plevel;
  std::vector<ProgImage> pimgs;
  int64_t prog_blocks = 0;
  std::vector<HuffTable> tables;
  std::vector<uint16_t> luts;    // kLutEntries per image
  std::vector<uint16_t> qtabs;   // natural order, 64 per table
  // Unstuffed segments, word aligned, 0xFF padded: nwords words at wptr.
  // When word_store is set, plan_batch asks it for the storage (pinned host
  // memory in the engine, so that the upload is one DMA); when it is unset or
  // returns nullptr, the words go to the `words` vector.
  std::function<uint32_t*(size_t)> word_store;
  std::vector<uint32_t> words;
  uint32_t* wptr = nullptr;
  size_t nwords = 0;
  // Device unstuffing (unstuff_device = 1): the host leaves the words to the
  // device (wptr stays null) and gives it instead every image's stuffed scan
  // bytes (nraw bytes at rptr, each scan 16-byte aligned; storage from
  // raw_store as for the words) and the starting state of every chunk of
  // kUnstuffChunk scan bytes.
  int unstuff_device = 0;
  std::function<uint8_t*(size_t)> raw_store;
  std::vector<uint8_t> raw;
  uint8_t* rptr = nullptr;
  size_t nraw = 0;
  std::vector<UnstuffChunk> chunks;
  int64_t total_blocks = 0, total_pixels = 0, scan_bytes = 0;
  int64_t entry_words = 0;       // entry buffer of the tiled decoder (regions, checkpoints, sinks)
  uint64_t out_bytes = 0;
};

// End of the ordered-redo launch of the tiled decoder that starts at work item
// a (of items [a, end)): the longest run [a, e) in which no item waits in the
// lookback for another item of the run -- each either starts its segment
// (t0 == 0) or follows an item before a.
inline uint32_t ordered_launch_end(const std::vector<WorkItem>& items, uint32_t a, uint32_t end) {
  uint32_t e = a + 1;
  while (e < end && !(items[e].t0 > 0 && items[e].prev_item >= a)) ++e;
  return e;
}

// src/decodeScanGPU.cu:11-24 on one scan (plan.cpp): the unstuffed segment
// bounds, the unstuffed segments written at dst + dst_off[k], and/or the
// chunk states of the device unstuffer; false on an out-of-sequence RSTn.
bool unstuff_segments(const uint8_t* src, size_t len, bool split, std::vector<size_t>* bounds, uint8_t* dst,
                      const size_t* dst_off, std::vector<UnstuffChunk>* chunks);

// Build the plan for n files; errors[i] gets each file's parse/stage result.
void plan_batch(Plan& p, int n, const uint8_t* const* files, const size_t* sizes, int* errors);
void build_table(const uint8_t bits[16], const uint8_t vals[256], HuffTable& t);
// First-level table l1 (kL1Size entries inside lut) of one Huffman table;
// second-level tables are taken from lut's pool (pool_used entries so far).
void build_lut(const uint8_t bits[16], const uint8_t vals[256], bool ac, uint16_t* lut, uint16_t* l1, int& pool_used);

}  // namespace cjd
