// Entropy-decode core of the MI355X baseline-JPEG decoder: the symbol decoder
// and the phases of the tile kernel, written as __host__ __device__ functions
// so that the same code runs in k_entropy (entropy.hip) and, lane by lane, in
// the CPU emulator used by the tests.
//
// Algorithm (replaces src/entropyRLEdecodeGPU.cu:7-238 and the host serial
// decode of src/decodeScanGPU.cu:435-451).  A restart segment is cut into
// tiles of S bits; a work item is a run of up to kWG tiles, one lane each.
// The decoder state at a block boundary is (bit, slot of the next block); a
// trajectory is the sequence of block boundaries a decoder started from some
// state passes through.  Huffman streams self-synchronise: trajectories
// started at different states soon share a boundary and are identical from
// there on.
//  1. Speculative pass: every tile is decoded from (its first bit, slot 0) to
//     its exit (the first block boundary at or after the tile end), writing
//     one entry per symbol and a checkpoint (state, entries so far) at its
//     first block boundary past every C bits.
//  2. Link: from its entry -- the exit of tile j-1 -- tile j is decoded into
//     a prefix region until the trajectory passes through one of the
//     speculative checkpoints; from there the tile is its speculative pass
//     (same exit, same entries).  A link that never joins decodes to the tile
//     end.  Tiles whose entry changes are linked again in fix-up rounds.
//  3. Counts (blocks, DC sums) of every tile come from its DC entries and are
//     prefix-summed.  Across items a decoupled lookback composes single-state
//     aggregates ("if the entry is E then the exit is X, plus counts") back
//     to a predecessor that published its true exit.
//  4. Every block on the true trajectory gets a record: where its AC entries
//     are and its absolute DC value (k_expand turns records into blocks).
//
// Symbol semantics follow src/decodeScanGPU.cu:28-100 bit for bit on valid
// streams.  An invalid code ends the block (an invalid DC code consumes one
// bit), so every trajectory advances.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "cjd_internal.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CJD_HD __host__ __device__
#else
#define CJD_HD
#endif

namespace cjd {

constexpr uint8_t kDeZigZag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                   12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                   35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                   58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

CJD_HD inline uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

// Sequential reader of an unstuffed segment, MSB first.  The next bits sit in
// a 64-bit accumulator fed one 32-bit word at a time from two resident 4-word
// chunks (current q, next n); a third chunk f is in flight.  Decode loops run
// in periods of kPeriod symbol steps: at the top of a period rotate() moves
// on to the next chunk when the current one is used up (q <- n, n <- f) and
// issues the load of the following one.  A step appends at most one word, and
// after rotate() at least 5 words are resident, so within a period no lane
// touches a register whose load may still be in flight: the only wait on
// global memory is for the chunk requested one period earlier.  Words are
// kept as loaded and byte-swapped when appended.  Words past the segment read
// as 0xFFFFFFFF (the reference pads with 0xFF beyond the end,
// src/decodeScanGPU.cu:32).
constexpr int kPeriod = 4;

struct BitReader {
  const uint32_t* w;
  uint32_t nwords;
  uint64_t acc;   // next bits, MSB-aligned; nacc of them valid (> 32 after refill)
  int nacc;
  uint32_t q0, q1, q2, q3;  // current chunk (as loaded)
  uint32_t n0, n1, n2, n3;  // next chunk
  uint32_t f0, f1, f2, f3;  // chunk after next, in flight
  uint32_t qi;    // next word to append: 0..3 in q, 4..7 in n
  uint32_t chunk; // word index of q0
  uint32_t pos;   // stream bit of the top of acc

  // Words past the data are 0xFF padding (>= kSegPadWords of them, enough for
  // the longest block a speculative decoder can read past the end plus the
  // prefetch); the chunk index is clamped so a read can never leave the
  // segment's buffer.
  CJD_HD inline void load(uint32_t wi, uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) const {
    wi = wi < nwords - 4 ? wi : nwords - 4;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 v = *reinterpret_cast<const uint4*>(w + wi);
    a = v.x;
    b = v.y;
    c = v.z;
    d = v.w;
#else
    a = w[wi];
    b = w[wi + 1];
    c = w[wi + 2];
    d = w[wi + 3];
#endif
  }
  CJD_HD inline uint32_t word(uint32_t i) const {
#if !defined(__HIP_DEVICE_COMPILE__)
    if (i > 7) __builtin_trap();  // CPU builds: the period invariant was broken
#endif
    const uint32_t a = (i & 4) ? n0 : q0, b = (i & 4) ? n1 : q1, c = (i & 4) ? n2 : q2, d = (i & 4) ? n3 : q3;
    const uint32_t lo = (i & 1) ? b : a, hi = (i & 1) ? d : c;
    return bswap32((i & 2) ? hi : lo);
  }
  CJD_HD inline void refill() {
    while (nacc <= 32) {
      acc |= (uint64_t)word(qi) << (32 - nacc);
      nacc += 32;
      ++qi;
    }
  }
  // Branch-free: the chunk shift is a select and the prefetch load is issued
  // every period (re-reading the same chunk when nothing moved), so the load
  // writes the loop-carried registers directly and nothing waits for it
  // before the next period.
  CJD_HD inline void rotate() {
    const bool r = qi >= 4;
    q0 = r ? n0 : q0;
    q1 = r ? n1 : q1;
    q2 = r ? n2 : q2;
    q3 = r ? n3 : q3;
    n0 = r ? f0 : n0;
    n1 = r ? f1 : n1;
    n2 = r ? f2 : n2;
    n3 = r ? f3 : n3;
    chunk += r ? 4u : 0u;
    qi -= r ? 4u : 0u;
    load(chunk + 8, f0, f1, f2, f3);
  }
  CJD_HD inline void seek(const uint32_t* words, uint32_t n, uint32_t p) {
    w = words;
    nwords = n;
    const uint32_t wi = p >> 5;
    chunk = wi & ~3u;
    load(chunk, q0, q1, q2, q3);
    load(chunk + 4, n0, n1, n2, n3);
    load(chunk + 8, f0, f1, f2, f3);
    qi = wi - chunk;
    acc = 0;
    nacc = 0;
    pos = p;
    refill();
    const int o = (int)(p & 31);
    acc <<= o;
    nacc -= o;
    refill();
    rotate();
  }
  CJD_HD inline uint32_t peek() const { return (uint32_t)(acc >> 32); }
  CJD_HD inline void skip(int n) {
    acc <<= n;
    nacc -= n;
    pos += (uint32_t)n;
    refill();
  }
};

// Huffman symbol of the code at the top of `bits`; len = code length, 0 for an
// invalid code.  Codes longer than the LUT: the length is found by comparing
// against all canonical limits at once (T.81 F.2.2.3).
CJD_HD inline int huff_sym(const HuffTable& T, uint32_t bits, int& len) {
  const uint32_t e = T.lut[bits >> (32 - kLutBits)];
  len = (int)(e & 31);
  if (len) return (int)((e >> 8) & 0xFF);
  const uint32_t v = bits >> 16;
  int l = kLutBits + 1;
#pragma unroll
  for (int k = kLutBits + 1; k <= 16; ++k) l += v >= T.lim[k] ? 1 : 0;
  if (l > 16) {
    len = 0;
    return 0;
  }
  len = l;
  return T.vals[(T.valoff[l] + (int)(v >> (16 - l))) & 0xFF];
}

// EXTEND (T.81 F.2.2.1) of the `size` bits following a `len`-bit code.
CJD_HD inline int extend_value(uint32_t bits, int len, int size) {
  if (!size) return 0;
  int v = (int)((bits << len) >> (32 - size));
  if (v < (1 << (size - 1))) v += (-1 << size) + 1;
  return v;
}

// Decode one block at the reader's position (src/decodeScanGPU.cu:78-100).
// err = 1 where the reference raises SYNTAX_ERROR; an invalid DC code
// consumes one bit so that speculative decoders always advance.  kWrite: AC
// coefficients into blk (natural order); the caller zeroes blk and stores DC.
template <bool kWrite>
CJD_HD inline void decode_block(BitReader& br, const HuffTable& dct, const HuffTable& act, int& dcdiff, int& err,
                                int16_t* blk) {
  br.rotate();
  uint32_t bits = br.peek();
  int len;
  int sym = huff_sym(dct, bits, len);
  if (!len) {
    err = 1;
    dcdiff = 0;
    br.skip(1);
    return;
  }
  int size = sym & 15;
  dcdiff = extend_value(bits, len, size);
  br.skip(len + size);
  int coef = 0;
  while (coef < 63) {
    br.rotate();  // at most one word is appended per symbol
    bits = br.peek();
    sym = huff_sym(act, bits, len);
    if (!len) {
      err = 1;
      return;
    }
    size = sym & 15;
    const int v = extend_value(bits, len, size);
    br.skip(len + size);
    if (!sym) break;  // EOB
    if (!size && sym != 0xF0) {
      err = 1;
      return;
    }
    coef += (sym >> 4) + 1;
    if (coef > 63) {
      err = 1;
      return;
    }
    if (kWrite) blk[kDeZigZag[coef]] = (int16_t)v;
  }
}

// Set status bits v in the word at p for launch `epoch`: a word of an older
// launch is replaced, one of this launch gains the bits (compare-and-swap, so
// that lanes of several workgroups can report the same image).
CJD_HD inline void status_or(uint32_t* p, uint32_t v, uint32_t epoch) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    const uint32_t nv = ((old >> 8) == epoch ? old : epoch << 8) | v;
    if (nv == old ||
        __hip_atomic_compare_exchange_strong(p, &old, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;
  }
#else
  uint32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
  for (;;) {
    const uint32_t nv = ((old >> 8) == epoch ? old : epoch << 8) | v;
    if (nv == old || __atomic_compare_exchange_n(p, &old, nv, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) break;
  }
#endif
}

// ------------------------------------------------------ tiled decoder --
//
// Lookup entry of a code of `len` bits for symbol `sym` (layout in
// cjd_internal.h).  AC symbols of size 0 other than EOB (0x00) and ZRL (0xF0)
// get 0: the reference raises SYNTAX_ERROR for them (src/decodeScanGPU.cu:91).
CJD_HD inline uint16_t lut_entry(int len, int sym, bool ac) {
  const int size = sym & 15;
  if (!ac) return (uint16_t)((len + size) | (size << 5));
  if (sym == 0x00) return (uint16_t)len;
  if (sym == 0xF0) return (uint16_t)(len | (16 << 9));
  if (!size) return 0;
  return (uint16_t)((len + size) | (size << 5) | (((sym >> 4) + 1) << 9));
}

constexpr int kRingRows = 9;  // 8 word rows + a copy of row 0

// Canonical decode of codes longer than kL1Bits bits of one table (T.81
// F.2.2.3): lim[l - 11] is the first left-justified 16-bit value past all
// codes of length <= l, valoff[l - 11] the index of the first symbol of
// length l minus the first code of length l.
struct CanonLong {
  uint32_t lim[6];
  int32_t valoff[6];
  uint8_t vals[256];
};

CJD_HD inline void canon_fill(CanonLong& d, const HuffTable& t) {
  for (int l = 11; l <= 16; ++l) {
    d.lim[l - 11] = t.lim[l];
    d.valoff[l - 11] = t.valoff[l];
  }
  for (int k = 0; k < 256; ++k) d.vals[k] = t.vals[k];
}

struct EntShared {
  uint16_t lut[kLutEntries];
  CanonLong canon[4];              // codes longer than kL1Bits when the second-level pool was full
  uint16_t tsel[2 * kMaxSlots];    // first-level base of the table for (slot, next symbol is DC)
  uint8_t dz[64];                  // zigzag index -> natural index
  uint32_t ring[kRingRows][kWG];   // each lane's window of its stream: word w in row w & 7
  // per local tile
  uint32_t sxp[kWG];               // speculative exit (bit)
  uint32_t xp[kWG];                // chain exit: assumed, then true
  uint32_t lp[kWG];                // entry the tile's link was decoded from
  uint8_t sxs[kWG], xs[kWG], ls[kWG];  // their slots
  uint8_t nck[kWG];                // usable checkpoints of the speculative pass
  uint8_t jslot[kWG];              // slot of the first suffix block
  uint16_t ss[kWG], sdn[kWG];      // speculative pass: steps, blocks (DC entries)
  uint16_t ps[kWG], pdn[kWG];      // link prefix: steps, blocks
  uint16_t js[kWG], jdn[kWG];      // speculative steps and blocks before the join
  union {
    struct {
      int32_t cnt[kWG][4];         // blocks / DC sums of components 0-2 of tile j
      int32_t tc[kWG][4];          // their inclusive prefix sums
    } c;
    struct {                       // link tasks (during the link and fix-up phases)
      uint32_t pos[kWG];
      uint32_t st[kWG];            // slot | (k + 1) << 4 | checkpoint << 11 | tile << 19
      uint32_t sd[kWG];            // step | DC entries << 16
      uint32_t last[kWG];          // DC-list entry of the current block
    } t;
  } u;
  int32_t ntask, nnext;
  uint32_t item;
  int32_t nneed;                   // tiles whose entry changed in the current fix-up round
  uint32_t lovf;                   // a link needed more entries than its tile has bits
  uint32_t look_ok;                // the lookback resolved the item's entry
  uint32_t entry_p, entry_s;       // true state at the exit of the lead-in tile
  int32_t entry_c[4];              //   and the segment's cumulative counts there
  int32_t wsum[kWG / 64][4];       // count scan: each wave's totals
};

// Constants of one work item.
// A segment's MCU structure: the image's slots for its interleaved scan, one
// slot (the component's) for a one-component scan.
CJD_HD inline int seg_nh(const SegmentDesc& sg, const ImageDesc& im) { return sg.ni ? 1 : im.nslots; }
CJD_HD inline int seg_comp(const SegmentDesc& sg, const ImageDesc& im, int slot) {
  return sg.ni ? sg.comp : im.slot_comp[slot];
}
CJD_HD inline int seg_dc(const SegmentDesc& sg, const ImageDesc& im, int slot) {
  return sg.ni ? sg.dc_id : im.slot_dc[slot];
}
CJD_HD inline int seg_ac(const SegmentDesc& sg, const ImageDesc& im, int slot) {
  return sg.ni ? sg.ac_id : im.slot_ac[slot];
}
// Coefficient-buffer (record) index of block a of the segment's scan (a
// counts from first_block): itself in an interleaved scan; for a
// one-component scan, where the interleaved scan holds that block.
CJD_HD inline int64_t seg_block(const SegmentDesc& sg, const ImageDesc& im, int64_t a) {
  if (!sg.ni) return a;
  const int64_t by = a / sg.bw, bx = a - by * sg.bw;
  return im.coef_base + ((by / sg.v) * im.mcus_x + bx / sg.h) * im.nslots + sg.slot0 + (by % sg.v) * sg.h +
         bx % sg.h;
}

struct EntCtx {
  const uint32_t* w;
  uint32_t nwords;
  const ImageDesc* im;
  const HuffTable* tabs;  // the image's canonical tables (codes the lookup tables could not hold)
  uint32_t* ebuf;
  SegmentDesc sg;
  WorkItem it;
  int nh;
  uint32_t S;
  int lead;     // 1 when local tile 0 is the lead-in (last tile of the previous item)
  int J;        // local tiles including the lead-in
  uint32_t tfirst;
  uint64_t slotinfo;  // 4 bits per MCU slot: component in bits 2-3

  CJD_HD void init(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                   const HuffTable* tables, uint32_t* entries, const WorkItem& item) {
    it = item;
    sg = segs[item.segment];
    im = images + sg.image;
    tabs = tables + (size_t)sg.tset * 4;
    ebuf = entries;
    nh = seg_nh(sg, *im);
    S = sg.tile_bits;
    w = words + sg.word_base;
    nwords = sg.nwords;
    lead = item.t0 > 0 ? 1 : 0;
    J = (int)item.ntiles + lead;
    tfirst = item.t0 - (uint32_t)lead;
    slotinfo = 0;
    for (int k = 0; k < nh; ++k) slotinfo |= (uint64_t)(seg_comp(sg, *im, k) << 2) << (4 * k);
  }
  CJD_HD int comp(int slot) const { return (int)((slotinfo >> (4 * slot + 2)) & 3); }
  CJD_HD uint32_t tile_end(uint32_t t) const { return (t + 1 >= sg.ntiles) ? sg.nbits : (t + 1) * S; }
  // Item layout (words from it.ebase), W = stream row (64, or J rounded up
  // to 16 below 64 tiles), nl = W * waves:
  //   speculative streams  nl * rs: entry of step s of local tile j at
  //                        ((j / 64) * rs + s - s % 4) * W + 4 (j % 64) + s % 4
  //                        (a wave's lanes write one contiguous row of 16-byte
  //                        groups per 4 steps)
  //   prefix streams       nl * rs, same layout
  //   speculative DC lists nl * rd (tile j at j * rd), prefix DC lists nl * rd
  //   checkpoints          nl * 2 * nck (tile j at j * 2 * nck)
  //   sink                 kSinkWords
  CJD_HD uint32_t W() const { return J >= 64 ? 64u : ((uint32_t)J + 15) & ~15u; }
  CJD_HD uint64_t nl() const { return (uint64_t)((J + 63) / 64) * W(); }
  CJD_HD uint32_t* sstream(int j) const {
    return ebuf + it.ebase + (uint64_t)(j >> 6) * sg.rs * W() + 4u * (uint32_t)(j & 63);
  }
  CJD_HD uint32_t* pstream(int j) const { return sstream(j) + nl() * sg.rs; }
  CJD_HD uint32_t* sdcl(int j) const { return ebuf + it.ebase + 2 * nl() * sg.rs + (uint64_t)j * sg.rd; }
  CJD_HD uint32_t* pdcl(int j) const { return sdcl(j) + nl() * sg.rd; }
  CJD_HD uint32_t* ckpts(int j) const {
    return ebuf + it.ebase + 2 * nl() * (sg.rs + sg.rd) + (uint64_t)j * 2 * sg.nck;
  }
  CJD_HD uint32_t* sink() const { return ebuf + it.ebase + 2 * nl() * (sg.rs + sg.rd + sg.nck); }
  // is local tile j the segment's first tile (decoded from its known start)?
  CJD_HD bool seg_first(int j) const { return tfirst + (uint32_t)j == 0; }
};

// Per-lane decoder of one trajectory.  The lane's next stream words sit in
// its column of sh.ring: word w in row w & 7 (byte-swapped to MSB-first), row
// 8 repeating row 0, so that any 32 bits at or after `pos` are two adjacent
// rows.  Words c .. c+7 are resident; words c+8 .. c+11 (f0-f3) are in
// flight.  Decoding runs in periods of kPeriod symbols: at the top of a
// period, once the lane is past word c+3, the in-flight chunk replaces words
// c .. c+3 and the next one is requested.  A symbol consumes at most 31 bits,
// so from pos < 32(c+4) the kPeriod symbols of a period read at most up to
// bit 32(c+4) - 1 + 31 (kPeriod - 1) + 31 < 32(c+8): inside the window.
//
// Every step writes one entry into the lane's stream at step index s, the
// same for all lanes of a wave, so that the wave writes one contiguous row:
// (natural index << 16 | value) of an AC coefficient; EOB and ZRL write a zero
// value at a position that no later coefficient of the block uses; the entry
// of a DC step is not used.  A DC symbol also appends (its step << 17 | value)
// to the lane's DC list, with bit 16 set when the block has an error (an
// error inside a block sets it in the block's list entry).
struct Lane {
  uint32_t pos, c;
  uint32_t f0, f1, f2, f3;
  uint32_t g0, g1, g2, g3;  // the load issued at the last period's top
  bool greal;               // ... and whether it was the lane's next chunk
  int k;         // -1 when the next symbol is a DC symbol, else the last AC index
  int slot;
  uint32_t s;    // step index
  uint32_t dn;   // DC-list entries
  uint32_t lastdc;
  uint64_t si, di;     // entry-buffer words of stream step 0 and DC-list entry 0
  uint32_t R, Rd;      // stream steps and DC-list entries available
  uint32_t W;          // stream row: words between steps
  bool ovf;      // a stream or list was full: writing into the item's sink
  bool hit;      // a capture happened in this period
  int m;         // next checkpoint
  uint32_t cap_p, cap_i;  // the capture: bit, slot | step << 4 | DC entries << 18
};


CJD_HD inline void ld4(const EntCtx& c, uint32_t wi, uint32_t& a, uint32_t& b, uint32_t& d, uint32_t& e) {
  wi = wi < c.nwords - 4 ? wi : c.nwords - 4;
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 v = *reinterpret_cast<const uint4*>(c.w + wi);
  a = v.x;
  b = v.y;
  d = v.z;
  e = v.w;
#else
  a = c.w[wi];
  b = c.w[wi + 1];
  d = c.w[wi + 2];
  e = c.w[wi + 3];
#endif
}

CJD_HD inline void ring_put(EntShared& sh, int tid, uint32_t row, uint32_t a, uint32_t b, uint32_t d, uint32_t e) {
  sh.ring[row][tid] = bswap32(a);
  sh.ring[row + 1][tid] = bswap32(b);
  sh.ring[row + 2][tid] = bswap32(d);
  sh.ring[row + 3][tid] = bswap32(e);
  sh.ring[row ? row : 8][tid] = bswap32(a);  // row 8 repeats row 0 (else row is written twice)
}

CJD_HD inline void lane_seek(EntShared& sh, int tid, const EntCtx& c, Lane& L, uint32_t p) {
  L.pos = p;
  L.c = (p >> 5) & ~3u;
  uint32_t a, b, d, e;
  ld4(c, L.c, a, b, d, e);
  ring_put(sh, tid, L.c & 7, a, b, d, e);
  ld4(c, L.c + 4, a, b, d, e);
  ring_put(sh, tid, (L.c + 4) & 7, a, b, d, e);
  ld4(c, L.c + 8, L.f0, L.f1, L.f2, L.f3);
  L.greal = false;
#if defined(__HIP_DEVICE_COMPILE__)
  // a seek is rare: wait here for the chunk in flight, so that the decode
  // loop's own prefetch only ever waits for the load of the previous period
  __builtin_amdgcn_s_waitcnt(0);
#endif
}

// Top of a period: the chunk requested at the last period's top becomes the
// one in flight, and a lane past word c+3 moves on (the in-flight chunk
// enters the ring) and requests its next chunk.  A lane moves about once in
// six periods; the load is issued by every lane every period all the same
// (a load under a branch made the compiler wait for it at the join), but a
// lane that did not move reads the segment's first words -- one cache line
// for the whole wave, where every lane re-reading its own chunk took a line
// per lane -- and drops them at the next period's top.  A lane moves at
// most once a period (a period consumes under 128 bits), so the chunk it
// requested is in place before it can be needed.
CJD_HD inline void lane_rotate(EntShared& sh, int tid, const EntCtx& c, Lane& L) {
  L.f0 = L.greal ? L.g0 : L.f0;
  L.f1 = L.greal ? L.g1 : L.f1;
  L.f2 = L.greal ? L.g2 : L.f2;
  L.f3 = L.greal ? L.g3 : L.f3;
  const bool mv = L.pos >= (L.c + 4) * 32;
  if (mv) {
    ring_put(sh, tid, L.c & 7, L.f0, L.f1, L.f2, L.f3);
    L.c += 4;
  }
  ld4(c, mv ? L.c + 8 : 0u, L.g0, L.g1, L.g2, L.g3);
  L.greal = mv;
}

// The 32 bits at `pos`: rows r, r+1 with r = word of bit pos-1, shifted so
// that bit pos comes first (at pos = 32w the shift is 0 and row r is unused).
CJD_HD inline uint32_t lane_peek(const EntShared& sh, int tid, uint32_t pos) {
  const uint32_t r = ((pos - 1) >> 5) & 7;
  const uint32_t hi = sh.ring[r][tid], lo = sh.ring[r + 1][tid];
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, (0u - pos) & 31);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> ((0u - pos) & 31));
#endif
}

// Entry of a code longer than kL1Bits: second-level table, or the canonical
// decode when the image's pool was full.
CJD_HD inline uint32_t lut_long(const EntShared& sh, uint32_t e, uint32_t bits, uint32_t tb) {
  if (e != kLutSlow) return sh.lut[(e & 0x7FFFu) + ((bits >> (32 - kL1Bits - kL2Bits)) & (kL2Size - 1))];
  const int t = (int)(tb >> kL1Bits);
  const CanonLong& T = sh.canon[t];
  const uint32_t v = bits >> 16;
  int l = kL1Bits + 1;
  for (int k = 0; k < 6; ++k) l += v >= T.lim[k] ? 1 : 0;
  if (l > 16) return 0u;
  const int sym = T.vals[(T.valoff[l - 11] + (int)(v >> (16 - l))) & 0xFF];
  return lut_entry(l, sym, t >= 2);
}

CJD_HD inline uint32_t ubfe(uint32_t v, uint32_t off, uint32_t width) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ubfe(v, off, width);
#else
  return width ? (v >> (off & 31)) & ((1u << width) - 1u) : 0u;
#endif
}

// One symbol (src/decodeScanGPU.cu:78-100 for one symbol: DC difference, or
// an AC run/size with EXTEND, EOB, ZRL, the coefficient index checks).
// Straight-line code returning the step's stream entry (the period stores its
// four entries as one group); the second-level lookup and the DC-list store
// are the only branches.  Before the symbol, the first block boundary at or
// after thr in this period is captured.
CJD_HD inline uint32_t lane_step(EntShared& sh, int tid, Lane& L, uint32_t thr, int nh, uint32_t* eb) {
  const uint32_t bits = lane_peek(sh, tid, L.pos);
  const bool isdc = L.k < 0;
  const bool hc = isdc && !L.hit && L.pos >= thr;
  const uint32_t info = (uint32_t)L.slot | (L.s << 4) | (L.dn << 18);
  L.cap_p = hc ? L.pos : L.cap_p;
  L.cap_i = hc ? info : L.cap_i;
  L.hit = L.hit || hc;
  const uint32_t tb = sh.tsel[(L.slot << 1) | (isdc ? 1 : 0)];
  uint32_t e = sh.lut[tb + (bits >> (32 - kL1Bits))];
  if (e & 0x8000u) e = lut_long(sh, e, bits, tb);
  const uint32_t adv = e & 31u, size = (e >> 5) & 15u, runp1 = (e >> 9) & 31u;
  const uint32_t raw = ubfe(bits, 32 - adv, size);
  const uint32_t t = (1u << size) - 1u;
  const int v = (int)raw - (raw <= (t >> 1) ? (int)t : 0);  // EXTEND (T.81 F.2.2.1)
  const bool invalid = adv == 0;
  L.pos += adv + ((invalid && isdc) ? 1u : 0u);  // an invalid DC code consumes one bit
  const int nk = isdc ? 0 : L.k + (int)runp1;
  const bool bad = invalid || nk > 63;
  const bool end = isdc ? invalid : (runp1 == 0 || nk >= 63);
  const uint32_t z = sh.dz[nk > 63 ? 63 : nk];
  const uint32_t nat = runp1 == 0 ? 63u : z;
  const uint32_t vv = (uint32_t)v & 0xFFFFu;
  const uint32_t entry = (nat << 16) | vv;
  const uint32_t dcent = (L.s << 17) | (invalid ? 0x10000u : 0u) | vv;
  if (isdc || bad) eb[L.di + (isdc ? L.dn : L.dn - 1)] = isdc ? dcent : (L.lastdc | 0x10000u);
  L.lastdc = isdc ? dcent : L.lastdc;
  L.dn += isdc ? 1u : 0u;
  L.s += 1;
  L.k = end ? -1 : nk;
  const int ns = L.slot + 1 == nh ? 0 : L.slot + 1;
  L.slot = end ? ns : L.slot;
  return entry;
}

// The four entries of a period, as one 16-byte store.
CJD_HD inline void store_group(uint32_t* p, uint32_t a, uint32_t b, uint32_t d, uint32_t e) {
#if defined(__HIP_DEVICE_COMPILE__)
  *reinterpret_cast<uint4*>(p) = make_uint4(a, b, d, e);
#else
  p[0] = a;
  p[1] = b;
  p[2] = d;
  p[3] = e;
#endif
}

// Checkpoint m of local tile j (2 words at ckpts(j) + 2m): the speculative
// trajectory's first block boundary at or after bit t*S + (m+1)*C, as (bit,
// slot | step << 4 | DC entries << 18 before that block).

// Return L to its capture (a block boundary).  The step index goes on: the
// stream entries written since the capture are marked unused.
CJD_HD inline void lane_restore(Lane& L, bool mark, uint32_t* eb) {
  const uint32_t sc = (L.cap_i >> 4) & 0x3FFFu;
  if (mark && !L.ovf)
    for (uint32_t q = sc; q < L.s; ++q) eb[L.si + (uint64_t)(q & ~3u) * L.W + (q & 3u)] = 0xFFFFFFFFu;
  L.pos = L.cap_p;
  L.slot = (int)(L.cap_i & 15u);
  L.dn = L.cap_i >> 18;
  L.k = -1;
}

// Decode a trajectory from L's state (a block boundary) up to its first block
// boundary at or after `end`, where L is left (step index = the capture's).
// The thresholds are the checkpoint bits t*S + (m+1)*C below `end`, then
// `end`; the steps of a period capture the first boundary past the current
// threshold and the top of the next period acts on it.  kLink = false
// (speculative pass): a capture below `end` is checkpoint m (not recorded
// after the streams filled; ncp = checkpoints recorded).  kLink = true: a
// capture below `end` is compared with the speculative pass's checkpoint m (m
// < nck) and the decode stops at the first one it equals (joined = m + 1).
// When a boundary past the next threshold may have passed in the same period,
// the lane decodes again from the capture (long blocks only).
static_assert(kPeriod == 4, "a period's four entries are one 16-byte group of the stream");

template <bool kLink>
CJD_HD inline bool lane_run(EntShared& sh, int tid, const EntCtx& c, Lane& L, uint32_t b0, uint32_t end,
                            uint32_t* ck, int nck, int& joined, int budget) {
  const uint32_t C = c.sg.ck_bits;
  const int M = (int)c.sg.nck;
  const uint32_t t1 = b0 + (uint32_t)(L.m + 1) * C;
  uint32_t thr = L.m < M && t1 < end ? t1 : end;
  uint32_t want_p = ~0u, want_i = ~0u;
  if (kLink && L.m < nck) {
    want_p = ck[2 * L.m];
    want_i = ck[2 * L.m + 1];
  }
  joined = 0;
  L.hit = false;
  L.cap_p = 0;
  L.cap_i = 0;
  for (int per = 0;; ++per) {
    if (L.hit) {
      L.hit = false;
      bool stop = L.cap_p >= end;
      if (!stop) {
        if (kLink) {
          if (L.m < nck && want_p == L.cap_p && ((want_i ^ L.cap_i) & 15u) == 0) {
            joined = L.m + 1;
            stop = true;
          }
        } else if (!L.ovf) {
          ck[2 * L.m] = L.cap_p;
          ck[2 * L.m + 1] = L.cap_i;
        }
      }
      if (stop) {
        lane_restore(L, false, c.ebuf);
        L.s = (L.cap_i >> 4) & 0x3FFFu;
        return true;
      }
      ++L.m;
      const uint32_t tn = b0 + (uint32_t)(L.m + 1) * C;
      thr = L.m < M && tn < end ? tn : end;
      if (kLink && L.m < nck) {
        want_p = ck[2 * L.m];
        want_i = ck[2 * L.m + 1];
      }
      if (L.pos >= thr) {
        lane_restore(L, true, c.ebuf);
        lane_seek(sh, tid, c, L, L.pos);
      }
    }
    if (budget >= 0 && per >= budget) return false;
    lane_rotate(sh, tid, c, L);
    if (L.ovf || L.s + kPeriod > L.R || L.dn + kPeriod + 1 > L.Rd) {
      // keep decoding for the exit, writing into the item's sink: the lane's
      // group of its 256-word stream part and entries 1-5 of the lane's
      // 8-entry list part (the word indexes wrap modulo 2^64 to land there)
      L.ovf = true;
      const uint64_t sk = (uint64_t)(c.sink() - c.ebuf);
      L.si = sk + 4u * (uint64_t)(tid & 63) - (uint64_t)L.s * L.W;
      L.di = sk + 4 * 64 + (uint64_t)(tid & 63) * 8 + 2 - (uint64_t)L.dn;
    }
    // a period starts at a step that is a multiple of 4: its entries are one
    // group of the stream
#if !defined(__HIP_DEVICE_COMPILE__)
    if (L.s & 3u) __builtin_trap();
#endif
    uint32_t* const grp = c.ebuf + (L.si + (uint64_t)L.s * L.W);
    const uint32_t e0 = lane_step(sh, tid, L, thr, c.nh, c.ebuf);
    const uint32_t e1 = lane_step(sh, tid, L, thr, c.nh, c.ebuf);
    const uint32_t e2 = lane_step(sh, tid, L, thr, c.nh, c.ebuf);
    const uint32_t e3 = lane_step(sh, tid, L, thr, c.nh, c.ebuf);
    store_group(grp, e0, e1, e2, e3);
  }
}

CJD_HD inline void lane_start(const EntCtx& c, Lane& L, const uint32_t* stream, const uint32_t* dcl, uint32_t R,
                              uint32_t Rd, int slot) {
  L.k = -1;
  L.slot = slot;
  L.s = 0;
  L.dn = 0;
  L.lastdc = 0;
  L.si = (uint64_t)(stream - c.ebuf);
  L.di = (uint64_t)(dcl - c.ebuf);
  L.W = c.W();
  L.m = 0;
  L.R = R;
  L.Rd = Rd;
  L.ovf = false;
}

// Phase 1: speculative pass of local tile tid from (tile start, slot 0) --
// the segment's known start for its first tile -- writing the tile's stream,
// DC list and checkpoints.
CJD_HD inline void ent_spec(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  const uint32_t t = c.tfirst + (uint32_t)tid;
  const uint32_t b0 = t * c.S;
  Lane L;
  lane_start(c, L, c.sstream(tid), c.sdcl(tid), c.sg.rs, c.sg.rd, 0);
  lane_seek(sh, tid, c, L, b0);
  int joined;
  lane_run<false>(sh, tid, c, L, b0, c.tile_end(t), c.ckpts(tid), 0, joined, -1);
  const int ncp = L.m < (int)c.sg.nck ? L.m : (int)c.sg.nck;
  sh.sxp[tid] = L.pos;
  sh.sxs[tid] = (uint8_t)L.slot;
  sh.xp[tid] = L.pos;
  sh.xs[tid] = (uint8_t)L.slot;
  sh.nck[tid] = (uint8_t)(L.ovf ? 0 : ncp);
  sh.ss[tid] = (uint16_t)(L.ovf ? 0 : L.s);
  sh.sdn[tid] = (uint16_t)(L.ovf ? 0 : L.dn);
  // until linked: the tile's blocks are all of its speculative pass
  sh.lp[tid] = b0;
  sh.ls[tid] = 0;
  sh.ps[tid] = 0;
  sh.pdn[tid] = 0;
  sh.js[tid] = 0;
  sh.jdn[tid] = 0;
  sh.jslot[tid] = 0;
}

// Links.  Local tile j linked from entry (p, s) decodes into its prefix
// stream until the trajectory passes through a checkpoint of the speculative
// pass (the rest of the tile is the speculative pass from there: same exit,
// its entries after the checkpoint) or reaches the tile end (a concrete
// exit).  An entry equal to the speculative pass's start joins at once.
// Links run as tasks in rounds of kLinkBudget periods; after each round the
// unfinished ones are packed into the lowest lanes, so that a few long links
// occupy one wave instead of every wave of the workgroup.
constexpr int kLinkBudget = 16;

// The join point's steps and DC entries are those of the speculative pass up
// to the checkpoint, so they cannot exceed the pass's totals; a checkpoint
// word that says otherwise (not what this pass wrote) leaves the tile without
// speculative blocks and the image to the serial decoder, so that no block
// count ever goes negative.
CJD_HD inline void set_lovf(EntShared& sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  sh.lovf = 1;
#else
  __atomic_store_n(&sh.lovf, 1u, __ATOMIC_RELAXED);
#endif
}

CJD_HD inline void link_joined(EntShared& sh, int j, uint32_t ps, uint32_t pdn, uint32_t info) {
  uint32_t js = (info >> 4) & 0x3FFF, jdn = info >> 18, jslot = info & 15;
  if (js > sh.ss[j] || jdn > sh.sdn[j]) {
    set_lovf(sh);
    js = sh.ss[j];
    jdn = sh.sdn[j];
    jslot = 0;
  }
  sh.ps[j] = (uint16_t)ps;
  sh.pdn[j] = (uint16_t)pdn;
  sh.js[j] = (uint16_t)js;
  sh.jdn[j] = (uint16_t)jdn;
  sh.jslot[j] = (uint8_t)jslot;
  sh.xp[j] = sh.sxp[j];
  sh.xs[j] = sh.sxs[j];
}

// Set tile j's link entry; false when it joins at once (no task needed).
CJD_HD inline bool link_begin(EntShared& sh, const EntCtx& c, int j, uint32_t p, uint32_t s) {
  const uint32_t b0 = (c.tfirst + (uint32_t)j) * c.S;
  sh.lp[j] = p;
  sh.ls[j] = (uint8_t)s;
  if (sh.sdn[j] != 0 && p == b0 && s == 0) {  // speculative entries kept and the same start
    link_joined(sh, j, 0, 0, 0);
    return false;
  }
  return true;
}

CJD_HD inline void task_put(EntShared& sh, int idx, int j, const Lane& L) {
  sh.u.t.pos[idx] = L.pos;
  sh.u.t.st[idx] = (uint32_t)L.slot | ((uint32_t)(L.k + 1) << 4) | ((uint32_t)L.m << 11) | ((uint32_t)j << 19);
  sh.u.t.sd[idx] = L.s | (L.dn << 16);
  sh.u.t.last[idx] = L.lastdc;
}

// A new task: tile j from its link entry.
CJD_HD inline void task_new(EntShared& sh, int idx, int j) {
  sh.u.t.pos[idx] = sh.lp[j];
  sh.u.t.st[idx] = (uint32_t)sh.ls[j] | ((uint32_t)j << 19);
  sh.u.t.sd[idx] = 0;
  sh.u.t.last[idx] = 0;
}

struct Task {
  uint32_t pos, st, sd, last;
};
CJD_HD inline Task task_get(const EntShared& sh, int idx) {
  return Task{sh.u.t.pos[idx], sh.u.t.st[idx], sh.u.t.sd[idx], sh.u.t.last[idx]};
}

// One round of task T on lane tid.  Returns true when the link finished
// (results written); else L holds the state to put back.
CJD_HD inline bool task_round(EntShared& sh, int tid, const EntCtx& c, const Task& T, Lane& L, int& j) {
  const uint32_t st = T.st, sd = T.sd;
  j = (int)(st >> 19);
  const uint32_t t = c.tfirst + (uint32_t)j;
  lane_start(c, L, c.pstream(j), c.pdcl(j), c.sg.rs, c.sg.rd, (int)(st & 15));
  L.k = (int)((st >> 4) & 127) - 1;
  L.m = (int)((st >> 11) & 255);
  L.s = sd & 0xFFFF;
  L.dn = sd >> 16;
  L.lastdc = T.last;
  lane_seek(sh, tid, c, L, T.pos);
  uint32_t* ck = c.ckpts(j);
  int joined;
  if (!lane_run<true>(sh, tid, c, L, t * c.S, c.tile_end(t), ck, sh.nck[j], joined, kLinkBudget)) return false;
  // under two bits per symbol over a whole tile: left to the serial decoder
  // (several lanes may set the flag)
  if (L.ovf) set_lovf(sh);
  if (joined) {
    link_joined(sh, j, L.ovf ? 0 : L.s, L.ovf ? 0 : L.dn, ck[2 * (joined - 1) + 1]);
  } else {
    sh.ps[j] = (uint16_t)(L.ovf ? 0 : L.s);
    sh.pdn[j] = (uint16_t)(L.ovf ? 0 : L.dn);
    sh.js[j] = sh.ss[j];
    sh.jdn[j] = sh.sdn[j];
    sh.jslot[j] = 0;
    sh.xp[j] = L.pos;
    sh.xs[j] = (uint8_t)L.slot;
  }
  return true;
}

// Phase 2: link entry of local tile tid -- the segment start for the
// segment's first tile, else the speculative exit of tile tid-1 (the lead-in
// has none); true when a link task is needed.
CJD_HD inline bool ent_phase2(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J || (c.lead && tid == 0)) return false;
  if (c.seg_first(tid)) return link_begin(sh, c, tid, 0, 0);
  return link_begin(sh, c, tid, sh.sxp[tid - 1], sh.sxs[tid - 1]);
}

// Fix-up round, read half (all lanes): was tile tid linked from an entry
// other than tile tid-1's current exit?  (p, s) = that exit.
CJD_HD inline bool ent_fix_need(const EntShared& sh, int tid, const EntCtx& c, uint32_t& p, uint32_t& s) {
  if (tid < 1 || tid >= c.J) return false;
  p = sh.xp[tid - 1];
  s = sh.xs[tid - 1];
  return p != sh.lp[tid] || s != sh.ls[tid];
}

CJD_HD inline void add_dc(int32_t* c3, int comp, int d) {
  c3[0] += comp == 0 ? d : 0;
  c3[1] += comp == 1 ? d : 0;
  c3[2] += comp == 2 ? d : 0;
}

// Blocks of local tile j on the true trajectory: its prefix blocks, then the
// speculative blocks from the join.  f(stream, dcl, d0, d1, send, slot) is
// called for the two runs: DC-list entries [d0, d1) of list dcl, whose
// blocks' AC entries are in the stream column `stream` (stride 64 words) and
// end at step send for the last block; slot is the first block's slot.
template <class F>
CJD_HD inline void tile_runs(const EntShared& sh, const EntCtx& c, int j, F&& f) {
  f(c.pstream(j), c.pdcl(j), 0u, (uint32_t)sh.pdn[j], (uint32_t)sh.ps[j], (int)sh.ls[j]);
  f(c.sstream(j), c.sdcl(j), (uint32_t)sh.jdn[j], (uint32_t)sh.sdn[j], (uint32_t)sh.ss[j], (int)sh.jslot[j]);
}

// f(e) for the DC-list entries dcl[d0 .. d1) in order, 16 entries per round
// trip to memory (lists are 16-byte aligned and followed by at least 16
// words of the item's region, so the reads past d1 stay inside it).
template <class F>
CJD_HD inline void dcl_for(const uint32_t* dcl, uint32_t d0, uint32_t d1, F&& f) {
  for (uint32_t g = d0 & ~15u; g < d1; g += 16) {
    uint32_t v[16];
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4* p = reinterpret_cast<const uint4*>(dcl + g);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 x = p[q];
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
#else
    for (int q = 0; q < 16; ++q) v[q] = dcl[g + q];
#endif
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (g + q >= d0 && g + q < d1) f(v[q]);
  }
}

// Counts of local tile tid (blocks, DC sums per component) from its DC
// entries; the lead-in counts nothing (its blocks belong to the previous item).
CJD_HD inline void ent_count(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  int32_t q[4] = {0, 0, 0, 0};
  if (!(c.lead && tid == 0)) {
    tile_runs(sh, c, tid, [&](const uint32_t*, const uint32_t* dcl, uint32_t d0, uint32_t d1, uint32_t, int slot) {
      dcl_for(dcl, d0, d1, [&](uint32_t e) {
        add_dc(q + 1, c.comp(slot), (int16_t)(e & 0xFFFFu));
        slot = slot + 1 == c.nh ? 0 : slot + 1;
      });
      q[0] += d1 > d0 ? (int32_t)(d1 - d0) : 0;
    });
  }
  for (int k = 0; k < 4; ++k) sh.u.c.cnt[tid][k] = q[k];
}

// Scan (all lanes; Hillis-Steele over kWG lanes in LDS with barriers between
// the steps supplied by the caller): tc[j] <- cnt[0] + ... + cnt[j].  A step
// with distance d first reads (ent_scan_read), then, after a barrier, adds.
CJD_HD inline void ent_scan_init(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  for (int q = 0; q < 4; ++q) sh.u.c.tc[tid][q] = sh.u.c.cnt[tid][q];
}
CJD_HD inline void ent_scan_read(const EntShared& sh, int tid, int d, int32_t* v) {
  for (int q = 0; q < 4; ++q) v[q] = tid >= d ? sh.u.c.tc[tid - d][q] : 0;
}
CJD_HD inline void ent_scan_write(EntShared& sh, int tid, const int32_t* v) {
  for (int q = 0; q < 4; ++q) sh.u.c.tc[tid][q] += v[q];
}

// Composition of predecessors' aggregates in the lookback.  An aggregate says
// "if the item's entry is (lead_p, lead_s), its exit is (agg_p, agg_s) and it
// adds agg_c"; absorbing an earlier item checks that its aggregate exit is
// the later item's assumed entry.
struct LookAcc {
  uint32_t lp, ls;     // assumed entry of the earliest absorbed item
  uint32_t xp, xs;     // exit of the latest
  int32_t c[4];
  bool empty, valid;
  CJD_HD void reset() {
    empty = true;
    valid = true;
    for (int q = 0; q < 4; ++q) c[q] = 0;
  }
  CJD_HD void absorb(const LookbackRec& r) {
    if (empty) {
      xp = r.agg_p;
      xs = r.agg_s;
      empty = false;
    } else {
      valid = valid && r.agg_p == lp && r.agg_s == ls;
    }
    lp = r.lead_p;
    ls = r.lead_s;
    for (int q = 0; q < 4; ++q) c[q] += r.agg_c[q];
  }
  // The exit of the latest absorbed item (or of r itself when none) from r's
  // inclusive state; false when an assumption on the way does not hold.
  CJD_HD bool apply(const LookbackRec& r, uint32_t& p, uint32_t& s, int32_t* cc) const {
    if (empty) {
      p = r.incl_p;
      s = r.incl_s;
      for (int q = 0; q < 4; ++q) cc[q] = r.incl_c[q];
      return true;
    }
    if (!valid || r.incl_p != lp || r.incl_s != ls) return false;
    p = xp;
    s = xs;
    for (int q = 0; q < 4; ++q) cc[q] = r.incl_c[q] + c[q];
    return true;
  }
};


// Lookback states of a record (the low 2 bits of its status word): an
// aggregate is published, the inclusive state is published, or the item
// failed (a wait gave up or a chain index left the segment's items; the
// item's image is marked unresolved and decoded again by the serial decoder).
constexpr int kLookAgg = 1, kLookIncl = 2, kLookFail = 3;

// The lookback walk of item `self`, whose segment's items are [first, self]
// and whose immediate predecessor is `prev`: compose the predecessors'
// aggregates back to one that has its inclusive state.  wait(q, incl)
// returns record q's status word once it holds a state of this launch (an
// inclusive or failed one when incl), or -1 when the wait gave up; load(q)
// returns record q's fields and is called only after wait(q) returned (on
// the device both read with agent-scope atomics: the records are written by
// work groups on other XCDs, each with an L2 of its own, and a plain load
// could hit a line that this XCD's L2 holds from before the write).  Every
// step moves to a lower item of the same segment, so a walk ends after at
// most self - first steps.  False: the entry cannot be resolved here (a wait
// gave up, a predecessor failed, or a chain index is out of range).
template <class Wait, class Load>
CJD_HD inline bool lookback_walk(const WorkItem* items, uint32_t first, uint32_t self, uint32_t prev, Wait&& wait,
                                 Load&& load, uint32_t& p, uint32_t& s, int32_t* cc) {
  LookAcc acc;
  acc.reset();
  if (prev < first || prev >= self) return false;
  uint32_t q = prev;
  for (;;) {
    const int st = wait(q, false);
    if (st < 0 || (st & 3) == kLookFail) return false;
    if ((st & 3) == kLookIncl) {
      if (acc.apply(load(q), p, s, cc)) return true;
      // an assumption on the way does not hold: take the immediate
      // predecessor's inclusive state
      const int sp = wait(prev, true);
      if (sp < 0 || (sp & 3) != kLookIncl) return false;
      acc.reset();
      return acc.apply(load(prev), p, s, cc);
    }
    acc.absorb(load(q));
    const uint32_t nq = items[q].prev_item;
    if (nq < first || nq >= q) return false;
    q = nq;
  }
}

// Words of a LookbackRec: the aggregate's [kRecAgg0, kRecAgg1), the
// inclusive state's [kRecIncl0, kRecIncl1); the status word is word 0.
constexpr int kRecWords = (int)(sizeof(LookbackRec) / 4);
constexpr int kRecAgg0 = 1, kRecAgg1 = 9, kRecIncl0 = 9, kRecIncl1 = 15;
static_assert(offsetof(LookbackRec, lead_p) == 4 * kRecAgg0 && offsetof(LookbackRec, incl_p) == 4 * kRecIncl0 &&
                  offsetof(LookbackRec, pad_) == 4 * kRecIncl1 && kRecWords >= kRecIncl1,
              "lookback record layout");

// Aggregate of this item (tc scanned: relative cumulative counts): assumes
// the item's entry is the speculative exit of its lead-in tile.
CJD_HD inline void ent_aggregate(const EntShared& sh, const EntCtx& c, LookbackRec& me) {
  me.lead_p = sh.sxp[0];
  me.lead_s = sh.sxs[0];
  me.agg_p = sh.xp[c.J - 1];
  me.agg_s = sh.xs[c.J - 1];
  for (int q = 0; q < 4; ++q) me.agg_c[q] = sh.u.c.tc[c.J - 1][q];
}

// Inclusive state of this item: the true exit of its last tile and the
// segment's cumulative counts there (tc absolute).
CJD_HD inline void ent_inclusive(const EntShared& sh, const EntCtx& c, LookbackRec& me) {
  me.incl_p = sh.xp[c.J - 1];
  me.incl_s = sh.xs[c.J - 1];
  for (int q = 0; q < 4; ++q) me.incl_c[q] = sh.u.c.tc[c.J - 1][q];
}

// After the lookback: whether the true entry differs from the assumed one
// (then the lead-in's exit is set to it and the fix-up rounds, the counts and
// the scan run again), and the absolute counts (every lane adds the entry's).
CJD_HD inline bool ent_entry_differs(const EntShared& sh, const EntCtx& c) {
  return c.lead && (sh.entry_p != sh.sxp[0] || sh.entry_s != sh.sxs[0]);
}
CJD_HD inline void ent_absolute(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J || !c.lead) return;
  for (int q = 0; q < 4; ++q) sh.u.c.tc[tid][q] += sh.entry_c[q];
}

// Final phase: the block records of local tile tid (absolute block index,
// absolute DC), the syntax errors of its blocks within the segment, and the
// truncation check at the segment's last tile.
CJD_HD inline void ent_records(const EntShared& sh, int tid, const EntCtx& c, BlockRec* recs, uint32_t* status,
                               uint32_t epoch) {
  if (tid >= c.J || (c.lead && tid == 0)) return;
  // only blocks of this segment get records (A in [first_block, seg_end)),
  // whatever the counts say
  const int64_t seg_end = c.sg.first_block + (int64_t)c.sg.nblocks;
  int64_t A = c.sg.first_block + (tid > 0 ? sh.u.c.tc[tid - 1][0] : 0);
  int32_t pred[3] = {0, 0, 0};
  if (tid > 0)
    for (int q = 0; q < 3; ++q) pred[q] = sh.u.c.tc[tid - 1][q + 1];
  bool err = false;
  tile_runs(sh, c, tid, [&](const uint32_t* stream, const uint32_t* dcl, uint32_t d0, uint32_t d1, uint32_t send,
                            int slot) {
    const uint32_t base = (uint32_t)(stream - c.ebuf);
    const uint32_t W = c.W();
    // block d's AC entries end where block d+1's DC step is: emit a block
    // when the next list entry is read
    auto emit = [&](uint32_t e, uint32_t sn) {
      const uint32_t st = e >> 17;
      const int comp = c.comp(slot);
      add_dc(pred, comp, (int16_t)(e & 0xFFFFu));
      if (A >= c.sg.first_block && A < seg_end) {
        BlockRec r;
        r.eidx = (uint32_t)entry_word(base, W, st + 1);
        r.nw = (uint16_t)((sn - st - 1) | ((W / 16 - 1) << 8));
        r.dc = (int16_t)(comp == 0 ? pred[0] : comp == 1 ? pred[1] : pred[2]);
        recs[seg_block(c.sg, *c.im, A)] = r;
        err = err || (e & 0x10000u);
      }
      ++A;
      slot = slot + 1 == c.nh ? 0 : slot + 1;
    };
    uint32_t prev = 0;
    bool have = false;
    dcl_for(dcl, d0, d1, [&](uint32_t e) {
      if (have) emit(prev, e >> 17);
      prev = e;
      have = true;
    });
    if (have) emit(prev, send);
  });
  const uint32_t sidx = (uint32_t)c.im->error_index;
  if (err) status_or(status + sidx, kStatusSyntax, epoch);
  if (sh.lovf) status_or(status + sidx, kStatusInconsistent, epoch);
  if (c.tfirst + (uint32_t)tid + 1 == c.sg.ntiles && sh.u.c.tc[tid][0] < (int32_t)c.sg.nblocks)
    status_or(status + sidx, kStatusTruncated, epoch);
}

// Serial decode of one whole segment from its known start (reference order,
// src/decodeScanGPU.cu:435-451).
CJD_HD inline void ent_serial_segment(const uint32_t* words, const ImageDesc* images, const SegmentDesc& sg,
                                      const HuffTable* tables, int16_t* coefs, BlockRec* recs, uint32_t* status,
                                      uint32_t epoch) {
  if (sg.prog) return;  // a progressive scan's interval: k_prog decodes it
  const ImageDesc& im = images[sg.image];
  const HuffTable* tab = tables + (size_t)sg.tset * 4;
  const int nh = seg_nh(sg, im);
  BitReader br;
  br.seek(words + sg.word_base, sg.nwords, 0);
  int slot = 0, dc[3] = {0, 0, 0};
  int16_t blk[64];
  for (uint32_t b = 0; b < sg.nblocks; ++b) {
    for (int k = 0; k < 64; ++k) blk[k] = 0;
    int d, err = 0;
    decode_block<true>(br, tab[seg_dc(sg, im, slot)], tab[seg_ac(sg, im, slot)], d, err, blk);
    if (err) {
      status_or(status + im.error_index, kStatusSyntax, epoch);
      return;
    }
    const int comp = seg_comp(sg, im, slot);
    dc[comp] += d;
    blk[0] = (int16_t)dc[comp];
    const int64_t gi = seg_block(sg, im, sg.first_block + b);
    int16_t* g = coefs + gi * 64;
    for (int k = 0; k < 64; ++k) g[k] = blk[k];
    BlockRec r;
    r.eidx = 0;
    r.nw = kRecDense;
    r.dc = blk[0];
    recs[gi] = r;
    if (++slot == nh) slot = 0;
  }
}


}  // namespace cjd
