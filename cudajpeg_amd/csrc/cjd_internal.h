// Internal structures shared by the host engine and the HIP kernels of the
// MI355X baseline-JPEG decoder.  Nothing here crosses the public C ABI.
#pragma once
#include <stdint.h>

namespace cjd {

constexpr int kLutBits = 9;                 // fast Huffman lookup width
constexpr int kLutSize = 1 << kLutBits;
constexpr int kMaxSlots = 10;               // blocks per MCU (T.81 B.2.3 limit)
constexpr int kWG = 256;                    // entropy workgroup size (4 wave64)
constexpr int kSegPadWords = 96;            // 0xFF words after every segment's data

// Tiled decoder lookup tables: per image kLutEntries uint16 -- four first-
// level tables of 2^kL1Bits entries (DC 0, DC 1, AC 0, AC 1), indexed by the
// next kL1Bits stream bits, then a pool of second-level tables of 2^kL2Bits
// entries for codes longer than kL1Bits, shared by the four.  Entry: bits 0-4
// adv (code length + extra bits), 5-8 size (extra bits), 9-13 run + 1 of an
// AC coefficient (16 for ZRL; 0 for EOB and DC symbols); 0 = no valid symbol.
// Bit 15 set: bits 0-14 index the second-level table of this prefix, or
// kLutSlow when the pool was full (canonical decode instead).
constexpr int kL1Bits = 10;
constexpr int kL1Size = 1 << kL1Bits;
constexpr int kL2Bits = 6;
constexpr int kL2Size = 1 << kL2Bits;
constexpr int kL2Pool = 2048;
constexpr int kLutEntries = 4 * kL1Size + kL2Pool;
constexpr uint16_t kLutSlow = 0xFFFF;
constexpr int kSinkWords = 4 * 64 + 64 * 8;  // per-item scratch of the tiled decoder (full streams write here)

// Device Huffman table: a 2^kLutBits direct table for codes up to kLutBits
// long plus canonical limits for longer codes (T.81 F.2.2.3).  LUT entry:
// bits 0-4 code length (0 = longer code), bits 8-15 the symbol.  lim[l] is the
// first left-justified 16-bit value past all codes of length <= l (non-
// decreasing in l), so the length of a long code is the first l with
// peek16 < lim[l]; its symbol is vals[valoff[l] + (peek16 >> (16 - l))].
struct HuffTable {
  uint32_t lut[kLutSize];
  uint32_t lim[18];         // lim[l] for l in 1..16
  int32_t valoff[18];       // index of the first symbol of length l minus the first code of length l
  uint8_t vals[256];
};

// Per-image constants used by both entropy and pixel kernels.
struct ImageDesc {
  int32_t width, height, ncomp;
  int32_t mcus_x, mcus_y;
  int32_t nslots;                 // blocks per MCU
  int32_t mcu_w, mcu_h;           // MCU size in pixels
  int32_t table_base;             // index of this image's 4 HuffTables
  int32_t qtab_base;              // index of this image's 4 quant tables (64 uint16, natural order)
  uint8_t slot_comp[kMaxSlots];   // component of each slot
  uint8_t slot_dc[kMaxSlots];     // DC table id (0..1) of each slot
  uint8_t slot_ac[kMaxSlots];     // AC table id (2..3) of each slot
  uint8_t comp_dq[3];             // quant table id per component
  uint8_t comp_sx[3], comp_sy[3]; // sampling factors
  uint8_t comp_xshift[3], comp_yshift[3]; // nearest-upsample shifts (CPUonly/src/pixelTransformCPU.c:110-127)
  uint8_t ntab;                   // distinct quant tables among the components (1..3)
  uint8_t comp_tix[3];            // index of each component's table among them
  uint8_t tab_dq[3];              // quant table id of each distinct table
  int32_t comp_first_slot[3];
  int32_t comp_stride[3];         // planar stride (src/format.cu:295)
  int32_t comp_plane_h[3];
  int64_t coef_base;              // first block of this image in the coefficient buffer
  int64_t nblocks;
  uint64_t out_offset;            // byte offset of the interleaved output in the output buffer
  uint64_t plane_offset[3];       // byte offsets of optional planar outputs (~0 = none)
  int32_t error_index;            // index into the per-image status array
  int32_t pad_;
};

// A segment is a run of the unstuffed entropy-coded stream whose starting
// decoder state is known: the start of the scan or the byte after an RSTn
// marker (DC predictors reset, first slot of an MCU).
struct SegmentDesc {
  uint64_t word_base;   // first 32-bit word of the segment in the scan buffer
  uint32_t nbits;       // data bits in the segment
  uint32_t nwords;      // readable words from word_base (data + 0xFF padding)
  uint32_t image;       // ImageDesc index
  uint32_t tile_bits;   // tile length S in bits
  uint32_t ntiles;
  uint32_t nblocks;     // blocks the segment must contain
  int64_t first_block;  // interleaved scan: global coefficient-buffer index of its first block;
                        // one-component scan: index of its first block in the scan
  uint32_t first_item;  // index of its first work item
  uint32_t ck_bits;     // checkpoint spacing C in bits
  uint32_t nck;         // checkpoints per tile (at C, 2C, ... < S)
  uint32_t rs;          // steps (entries) of a tile's speculative or prefix stream
  uint32_t rd;          // entries of a tile's DC list (per stream)
  uint32_t tset;        // Huffman table set of its scan: HuffTables 4 tset .. 4 tset + 3, lookup table tset
  // A one-component scan (CJD_EXT_NONINTERLEAVED, ni = 1) codes component
  // comp's own block grid, bw blocks a row, in raster order (T.81 A.2.2), one
  // block per MCU of the scan; its block (bx, by) is the interleaved layout's
  // slot slot0 + (by % v) h + bx % h of MCU (bx / h, by / v).
  uint8_t ni, comp, h, v;
  uint8_t dc_id, ac_id, slot0;  // the scan's DC (0..1) and AC (2..3) table ids
  uint8_t prog;                 // a progressive scan's interval: decoded by k_prog, no entropy work items
  uint32_t bw;
};

// One lane of the progressive decoder (CJD_EXT_PROGRESSIVE): one restart
// interval of one scan (T.81 G.1.2).  A scan of several components runs over
// the frame's MCUs, bpu blocks each (each component's h x v blocks in frame
// order, as the interleaved scan); a one-component scan over the component's
// own grid of bw blocks a row, one block per unit (T.81 A.2.2).
struct ProgTask {
  uint32_t seg;      // SegmentDesc of its unstuffed words, image and Huffman table set
  uint32_t u0;       // first unit (MCU, or block of a one-component scan) of the interval
  uint32_t nunits;
  uint32_t bw;       // one-component scans: blocks a row of the component's grid
  uint8_t ncomp, ss, se, ah;  // components, spectral band Ss..Se (zigzag), successive approximation
  uint8_t al, bpu, h, v;      // h, v: a one-component scan's sampling factors
  uint8_t comps[3], last;     // components (frame indexes); last: the scan's final interval
  uint8_t dcs[3], ac;         // DC table ids (0..1) of the components, the AC table id (2..3)
  // a scan of several components: the MCU's blocks in decode order, each
  // block's slot (4 bits each) and its component's index in the scan (2 bits each)
  uint64_t slots;
  uint32_t cis, pad_;
};

// A progressive image's coefficient blocks: [coef_base, coef_base + nblocks)
// of the coefficient buffer, block0 = the blocks of the progressive images
// before it (k_prog_prepare clears them and writes their records).
struct ProgImage {
  int64_t coef_base, nblocks, block0;
};

// One entropy workgroup: tiles [t0, t0 + ntiles) of a segment.  When t0 > 0 the
// workgroup also decodes tile t0-1 (the lead-in) to link to its predecessor.
struct WorkItem {
  uint32_t segment;
  uint32_t t0;
  uint32_t ntiles;
  uint32_t prev_item;   // predecessor in the same segment, ~0 for the first
  uint64_t ebase;       // entry-buffer word of the item's streams, DC lists and checkpoints
};

// Stream row of a work item of J local tiles: W lanes (W = 64, or J rounded
// up to 16 for J < 64); a stream row holds 4 consecutive steps of each of
// its W lanes (4 W words, lane-major: a lane's 4 entries are one 16-byte
// group), so a wave writes one contiguous 1 KB row per 4 steps; nl = lanes
// over all waves.
inline uint32_t item_row(uint32_t J) { return J >= 64 ? 64u : (J + 15) & ~15u; }
inline uint64_t item_lanes(uint32_t J) { return (uint64_t)((J + 63) / 64) * item_row(J); }
// Entry-buffer words of a work item (layout in entropy_core.h).
inline uint64_t item_words(uint32_t J, uint32_t rs, uint32_t rd, uint32_t nck) {
  return item_lanes(J) * (2ull * rs + 2ull * rd + 2ull * nck) + kSinkWords;
}

// Word of entry i of a stream whose entry 0 is at word e (a lane's stream
// of row width W): entries go 4 to a 16-byte group, groups 4 W words apart.
inline constexpr uint64_t entry_word(uint64_t e, uint32_t W, uint32_t i) {
  return (e & ~3ull) + (uint64_t)(((uint32_t)(e & 3u) + i) >> 2) * 4u * W + (((uint32_t)e + i) & 3u);
}

// Where the quantised coefficients of one block are, after the entropy
// decode: AC entries ebuf[entry_word(eidx, W, i)] for i < n (bits 16-21
// natural index, 0-15 value; 0xFFFFFFFF = none) with n in bits 0-7 of nw and
// the stream row W = 16 (1 + bits 8-9 of nw), and the absolute DC value.
struct BlockRec {
  uint32_t eidx;
  uint16_t nw;
  int16_t dc;
};
// nw flag: the block is in the dense coefficient buffer (written by the
// serial decoder), not in entries.
constexpr uint16_t kRecDense = 0x8000;

// Lookback record of one work item.  The aggregate holds under one
// assumption: if the item's entry (the exit of its lead-in tile) is the
// speculative exit (lead_p, lead_s) of that tile, the exit of its last tile is
// (agg_p, agg_s) and the item adds agg_c (blocks, DC sums of components 0-2).
// The inclusive state is the true exit of the last tile with the segment's
// cumulative counts there.
//
// Epochs.  Every entropy launch on a batch's scratch carries an epoch number
// (1 .. kEpochMax, one more than the launch before), and the scratch words
// are tagged with the epoch of the launch that wrote them, so nothing has to
// clear them between decodes: a record or status word of an older launch
// simply does not match.  The status word of a record is
// lookback_word(epoch, 1) once its aggregate is published and
// lookback_word(epoch, 2) with its inclusive state (written last), or
// lookback_word(epoch, 3) when the item gave up (a wait ran out of time or
// the chain was broken): its successors give up too, and the image is decoded
// again by the serial decoder.
struct LookbackRec {
  int32_t status;         // lookback_word(epoch, 1 aggregate | 2 inclusive | 3 failed)
  uint32_t lead_p, lead_s;
  uint32_t agg_p, agg_s;
  int32_t agg_c[4];
  uint32_t incl_p, incl_s;
  int32_t incl_c[4];
  int32_t pad_[3];  // 72 bytes
};

// Epochs of direct launches run 1 .. kEpochMax; kGraphEpoch is the fixed
// epoch of every entropy launch of a captured decode (whose first kernel
// clears the counters, status words and record states instead).
constexpr uint32_t kEpochMax = 0xFFFFFEu;
constexpr uint32_t kGraphEpoch = 0xFFFFFFu;

// Head of a batch's scratch: for each kind j of entropy launch (j < 4: run j
// of a decode in parts; kRedoCounter: one image decoded again alone), a 64-bit
// work-item ticket counter (words 2j, 2j+1) and a 64-bit counter of work items
// finished (words 2 kLaunchCounters + 2j, + 1).  Neither is ever cleared
// between decodes: launch kind j always takes its tickets from counter j, so
// launches never share one, and after the decodes both counters of kind j
// equal the tickets the host handed out to it.
constexpr int kLaunchCounters = 8;
constexpr int kRedoCounter = 4;
constexpr int kScratchHead = 4 * kLaunchCounters;

// Default limit of one lookback wait, in ticks of the 100 MHz real-time
// counter (200 ms: far above the longest wait of a valid decode).
constexpr uint64_t kDefaultWaitTicks = 20000000ull;
inline constexpr int32_t lookback_word(uint32_t epoch, int st) { return (int32_t)(epoch << 2 | (uint32_t)st); }

// Per-image status bits.  A status word holds (epoch << 8) | bits; a word of
// another epoch holds no bits for this one (see status_or in entropy_core.h).
inline constexpr uint32_t status_bits(uint32_t word, uint32_t epoch) {
  return (word >> 8) == epoch ? word & 0xFFu : 0u;
}
enum : uint32_t {
  kStatusSyntax = 1u,        // invalid Huffman code / coefficient index (reference SYNTAX_ERROR)
  kStatusTruncated = 2u,     // scan ended before all blocks were decoded
  kStatusUnresolved = 4u,    // the decoder state could not be resolved; rerun with the serial decoder
  kStatusInconsistent = 8u,  // the tiled decoder could not hold the image's entries (invalid stream); rerun serially
  kStatusFallback = 0x100u,  // host: the image was decoded again by the serial decoder
  kStatusRelaunch = 0x200u,  // host: the image was decoded again by the tiled decoder alone
  kStatusOrdered = 0x400u,   // host: ... by the tiled decoder in ordered launches (no lookback wait)
};
// host: bits 16-23 of an image's status after a redo hold the first tiled
// launch's bits that led to it
constexpr int kStatusFirstShift = 16;

enum IdctMode : int { kIdctInt = 0, kIdctFloat = 1, kIdctMfma = 2 };
// 16-byte MFMA table operands per image: 3 tables x 2 K halves x (hi, lo) x 256 lanes.
constexpr int kMfmaTabVecs = 3 * 2 * 2 * 256;

// One pixel workgroup: `count` consecutive MCUs of one MCU row of an image.
struct PixItem {
  uint32_t image;
  uint32_t mcu_first;
  uint32_t count;
  uint32_t pad_;
};
constexpr int kPixBlocks = 128;  // max 8x8 blocks per pixel workgroup

}  // namespace cjd
