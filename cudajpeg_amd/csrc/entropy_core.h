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
//  1. Main trajectory of each tile: from the tile's first bit one trajectory
//     per MCU slot h ("hypothesis" h: a block of slot h starts there) is
//     decoded until its first boundary at or after the sync point (D bits into
//     the tile).  The boundary most of them reach is the tile's D-state; the
//     main trajectory continues from it to the tile's exit (its first
//     boundary at or after the tile end), counting blocks and DC sums.
//  2. Link: from the main exit of tile j-1 a lane decodes into tile j up to
//     its first boundary at or after the sync point.  If that is the D-state,
//     the link has joined the main trajectory: tile j's exit is its main exit.
//     Otherwise the lane decodes on to the tile end (a concrete exit).
//  3. Chain: assuming the item's entry is the main exit of its lead-in tile
//     (the previous item's last tile), every tile's exit follows from the
//     links; a tile whose link did not start from its true entry is decoded
//     again from that entry (rare).  Counts are prefix-summed.  Across items a
//     decoupled lookback composes these single-state aggregates ("if the
//     entry is E then the exit is X, plus counts") back to a predecessor that
//     published its true exit.
//  4. Every tile is decoded once more from its true entry (bit, slot, block
//     index, DC predictors), writing quantised coefficients, and its exit is
//     checked against the chain.
//
// Symbol semantics follow src/decodeScanGPU.cu:28-100 bit for bit on valid
// streams.  During speculation an invalid code ends the block (an invalid DC
// code consumes one bit), so every trajectory advances.
#pragma once
#include <stdint.h>
#include <string.h>

#include "cjd_internal.h"

#if defined(__HIPCC__)
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

CJD_HD inline void status_or(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicOr(p, v);
#else
  *p |= v;
#endif
}

// ------------------------------------------------------------ LDS state --
// Decoder state at a block boundary packed as bit << 4 | slot.
CJD_HD inline uint64_t pack_state(uint32_t p, uint32_t s) { return ((uint64_t)p << 4) | s; }

struct EntShared {
  HuffTable tab[4];
  uint8_t dz[64];                 // zigzag index -> natural index
  uint32_t ep[kWG], es[kWG];      // main exit of each local tile (bit, slot)
  uint32_t tp[kWG], ts[kWG];      // chain: exit of tile j (assumed, then true)
  int32_t tc[kWG][4];             // chain: cumulative blocks / DC sums at the exit of tile j
                                  //   (relative to the item entry, then absolute)
  union {
    int16_t coef[kWG][64];        // final decode: one block per lane
    struct {
      uint64_t hyp[kMaxSlots][kWG];  // hypothesis states at a sync point
      uint64_t dstate[kWG];          // the tile's D-state (main trajectory)
      uint32_t dthr[kWG];            // the sync point it refers to
      uint32_t up[kWG], us[kWG];     // chain: entry tile j's exit was computed from
      int32_t mc[kWG][4];            // main counts from the D-state to the exit
      int32_t cnt[kWG][4];           // chain: blocks / DC sums gained in tile j
    } a;
  } u;
  uint32_t item;
  int32_t nchg;                   // tiles relinked in the current fix-up round
  uint32_t entry_p, entry_s;      // true state at the exit of the lead-in tile
  int32_t entry_c[4];             //   and the segment's cumulative counts there
};

// Constants of one work item.
struct EntCtx {
  const uint32_t* w;
  uint32_t nwords;
  const ImageDesc* im;
  SegmentDesc sg;
  WorkItem it;
  int nh;       // hypotheses = blocks per MCU
  uint32_t S;   // tile bits
  uint32_t D1;  // first sync point, bits into a tile
  uint32_t D;   // second sync point
  int lead;     // 1 when local tile 0 is the lead-in (last tile of the previous item)
  int J;        // local tiles including the lead-in
  uint32_t tfirst;
  uint64_t slotinfo;  // 4 bits per MCU slot: DC table (bit 0), AC table - 2 (bit 1), component (bits 2-3)

  CJD_HD void init(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs, const WorkItem& item) {
    it = item;
    sg = segs[item.segment];
    im = images + sg.image;
    nh = im->nslots;
    S = sg.tile_bits;
    D1 = sg.sync1_bits;
    D = sg.sync_bits;
    w = words + sg.word_base;
    nwords = sg.nwords;
    lead = item.t0 > 0 ? 1 : 0;
    J = (int)item.ntiles + lead;
    tfirst = item.t0 - (uint32_t)lead;
    slotinfo = 0;
    for (int k = 0; k < nh; ++k)
      slotinfo |= (uint64_t)((im->slot_dc[k] & 1) | ((im->slot_ac[k] & 1) << 1) | (im->slot_comp[k] << 2))
                  << (4 * k);
  }
  CJD_HD int comp(int slot) const { return (int)((slotinfo >> (4 * slot + 2)) & 3); }
  CJD_HD uint32_t tile_end(uint32_t t) const { return (t + 1 >= sg.ntiles) ? sg.nbits : (t + 1) * S; }
  CJD_HD uint32_t sync_point(uint32_t t, uint32_t d) const {
    const uint32_t e = tile_end(t), q = t * S + d;
    return q < e ? q : e;
  }
};

CJD_HD inline void add_dc(int32_t* c3, int comp, int d) {
  c3[0] += comp == 0 ? d : 0;
  c3[1] += comp == 1 ? d : 0;
  c3[2] += comp == 2 ? d : 0;
}

// One Huffman symbol of the block-structured stream (the body of
// decode_block, one symbol per call, so that the lanes of a wave stay in step
// symbol by symbol instead of waiting for each other's longest block).
// State: slot of the current block; coef = -1 when the next symbol is a DC
// (block start), else the index of the last coefficient.  Out: DC difference
// (0 for AC symbols), natural position and value of a coefficient to store
// (wpos = -1 for none).  Returns true when the block ended with this symbol.
CJD_HD inline bool sym_step(BitReader& br, const EntShared& sh, uint64_t slotinfo, int nh, int& slot, int& coef,
                            int& dcdiff, int& wpos, int& wval, int& err) {
  const uint32_t bits = br.peek();
  const bool isdc = coef < 0;
  const int sinfo = (int)((slotinfo >> (4 * slot)) & 15);
  const int tidx = isdc ? (sinfo & 1) : 2 + ((sinfo >> 1) & 1);
  int len;
  const int sym = huff_sym(sh.tab[tidx], bits, len);
  const int size = sym & 15;
  const int v = extend_value(bits, len, size);
  // an invalid DC code consumes one bit, an invalid AC code none
  br.skip(len ? len + size : (isdc ? 1 : 0));
  const int nc = coef + (sym >> 4) + 1;
  const bool bad = !len || (!isdc && sym && !size && sym != 0xF0) || (!isdc && sym && nc > 63);
  const bool end = isdc ? !len : (bad || !sym || nc >= 63);
  err |= bad ? 1 : 0;
  dcdiff = isdc ? v : 0;
  wpos = (!isdc && !bad && sym) ? (int)sh.dz[nc] : -1;
  wval = v;
  coef = isdc ? 0 : nc;
  if (end) {
    coef = -1;
    if (++slot == nh) slot = 0;
  }
  return end;
}

// The hypothesis state most hypotheses of the lane's tile share (ties: the
// lowest hypothesis); n = how many share it.
CJD_HD inline uint64_t ent_consensus(const EntShared& sh, int tid, int nh, int& n) {
  uint64_t best = sh.u.a.hyp[0][tid];
  int nbest = 0;
  for (int h = 0; h < nh; ++h) {
    const uint64_t v = sh.u.a.hyp[h][tid];
    int k = 0;
    for (int g = 0; g < nh; ++g) k += sh.u.a.hyp[g][tid] == v ? 1 : 0;
    if (k > nbest) {
      nbest = k;
      best = v;
    }
  }
  n = nbest;
  return best;
}

// Phase 1: main trajectory of local tile tid.  One loop, one symbol per step:
//   stage 1: hypotheses 0..nh-1 in turn from the tile start to the first sync
//            point; if they all agree there, that state is the D-state;
//   stage 2: otherwise each hypothesis continues from its stage-1 state to the
//            second sync point and the consensus there is the D-state;
//   main:    from the D-state to the tile end, counting blocks and DC sums.
// Tile 0 of a segment starts from the known state (bit 0, slot 0): its main
// trajectory is the true one from the start.
CJD_HD inline void ent_phase1(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  const int nh = c.nh, kMain = 2 * nh;
  const uint32_t t = c.tfirst + (uint32_t)tid;
  const uint32_t b0 = t * c.S, end = c.tile_end(t), thr1 = c.sync_point(t, c.D1), thr2 = c.sync_point(t, c.D);
  BitReader br;
  br.seek(c.w, c.nwords, b0);
  int mode = t == 0 ? kMain : 0;
  int slot = 0, coef = -1, err = 0;
  int32_t nb = 0, dc[3] = {0, 0, 0};
  sh.u.a.dstate[tid] = ~0ull;
  sh.u.a.dthr[tid] = end;
  for (bool done = false; !done;) {
    br.rotate();
    for (int u = 0; u < kPeriod; ++u) {
      if (coef < 0) {
        const uint32_t p = br.pos;
        if (mode < kMain && p >= (mode < nh ? thr1 : thr2)) {
          const int h = mode < nh ? mode : mode - nh;
          sh.u.a.hyp[h][tid] = pack_state(p, (uint32_t)slot);
          ++mode;
          uint64_t next = 0;  // state to continue from
          if (mode < nh) {
            next = pack_state(b0, (uint32_t)mode);
          } else if (mode == nh) {
            int n;
            const uint64_t m = ent_consensus(sh, tid, nh, n);
            if (n == nh || thr1 >= thr2) {  // unanimous, or no second sync point
              mode = kMain;
              sh.u.a.dstate[tid] = m;
              sh.u.a.dthr[tid] = thr1;
              next = m;
            } else {
              next = sh.u.a.hyp[0][tid];
            }
          } else if (mode < kMain) {
            next = sh.u.a.hyp[mode - nh][tid];
          } else {
            int n;
            const uint64_t m = ent_consensus(sh, tid, nh, n);
            sh.u.a.dstate[tid] = m;
            sh.u.a.dthr[tid] = thr2;
            next = m;
          }
          br.seek(c.w, c.nwords, (uint32_t)(next >> 4));
          slot = (int)(next & 15);
          continue;
        }
        if (mode == kMain && p >= end) {
          done = true;
          break;
        }
        if (mode == kMain) ++nb;
      }
      const int s0 = slot;
      int d, wp, wv;
      sym_step(br, sh, c.slotinfo, nh, slot, coef, d, wp, wv, err);
      if (mode == kMain) add_dc(dc, c.comp(s0), d);
    }
  }
  sh.ep[tid] = br.pos;
  sh.es[tid] = (uint32_t)slot;
  sh.u.a.mc[tid][0] = nb;
  sh.u.a.mc[tid][1] = dc[0];
  sh.u.a.mc[tid][2] = dc[1];
  sh.u.a.mc[tid][3] = dc[2];
}

// Decode local tile j from the state (p, slot) at its entry: up to the first
// boundary at or after the tile's D-state sync point; if that is the D-state,
// the rest is the main trajectory (its exit and counts are taken from phase
// 1), otherwise the decode goes on to the tile end.  (p, slot) become the
// exit, cnt the blocks / DC sums gained in tile j.
CJD_HD inline void ent_relink(const EntShared& sh, const EntCtx& c, int j, uint32_t& p, int& slot, int32_t* cnt) {
  const uint32_t t = c.tfirst + (uint32_t)j;
  const uint32_t end = c.tile_end(t), thr = sh.u.a.dthr[j];
  const uint64_t ds = sh.u.a.dstate[j];
  int32_t nb = 0, dc[3] = {0, 0, 0};
  bool checked = false;
  BitReader br;
  br.seek(c.w, c.nwords, p);
  int coef = -1, err = 0;
  for (bool done = false; !done;) {
    br.rotate();
    for (int u = 0; u < kPeriod; ++u) {
      if (coef < 0) {
        const uint32_t q = br.pos;
        if (!checked && q >= thr) {
          checked = true;
          if (pack_state(q, (uint32_t)slot) == ds) {
            cnt[0] = nb + sh.u.a.mc[j][0];
            cnt[1] = dc[0] + sh.u.a.mc[j][1];
            cnt[2] = dc[1] + sh.u.a.mc[j][2];
            cnt[3] = dc[2] + sh.u.a.mc[j][3];
            p = sh.ep[j];
            slot = (int)sh.es[j];
            return;
          }
        }
        if (q >= end) {
          done = true;
          break;
        }
        ++nb;
      }
      const int s0 = slot;
      int d, wp, wv;
      sym_step(br, sh, c.slotinfo, c.nh, slot, coef, d, wp, wv, err);
      add_dc(dc, c.comp(s0), d);
    }
  }
  p = br.pos;
  cnt[0] = nb;
  cnt[1] = dc[0];
  cnt[2] = dc[1];
  cnt[3] = dc[2];
}

// Phase 2: link local tile tid (>= 1) from the main exit of tile tid-1, and
// seed the chain with the assumed states (tile 0: its main exit; its counts
// belong to this item only when it is the segment's first tile).
CJD_HD inline void ent_phase2(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  if (tid == 0) {
    sh.tp[0] = sh.ep[0];
    sh.ts[0] = sh.es[0];
    for (int q = 0; q < 4; ++q) sh.u.a.cnt[0][q] = c.lead ? 0 : sh.u.a.mc[0][q];
    return;
  }
  uint32_t p = sh.ep[tid - 1];
  int slot = (int)sh.es[tid - 1];
  int32_t cnt[4];
  sh.u.a.up[tid] = p;
  sh.u.a.us[tid] = (uint32_t)slot;
  ent_relink(sh, c, tid, p, slot, cnt);
  sh.tp[tid] = p;
  sh.ts[tid] = (uint32_t)slot;
  for (int q = 0; q < 4; ++q) sh.u.a.cnt[tid][q] = cnt[q];
}

// Fix-up round, read half (all lanes): does tile tid's exit come from an
// entry other than tile tid-1's current exit?  (p, s) = that exit.
CJD_HD inline bool ent_fix_need(const EntShared& sh, int tid, const EntCtx& c, uint32_t& p, uint32_t& s) {
  if (tid < 1 || tid >= c.J) return false;
  p = sh.tp[tid - 1];
  s = sh.ts[tid - 1];
  return p != sh.u.a.up[tid] || s != sh.u.a.us[tid];
}
// Fix-up round, write half: decode tile tid again from (p, s).  Rounds repeat
// until no tile changes; a run of k consecutive tiles off the main trajectory
// takes k rounds.
CJD_HD inline void ent_fix_apply(EntShared& sh, int tid, const EntCtx& c, uint32_t p, uint32_t s) {
  sh.u.a.up[tid] = p;
  sh.u.a.us[tid] = s;
  int slot = (int)s;
  int32_t cnt[4];
  ent_relink(sh, c, tid, p, slot, cnt);
  sh.tp[tid] = p;
  sh.ts[tid] = (uint32_t)slot;
  for (int q = 0; q < 4; ++q) sh.u.a.cnt[tid][q] = cnt[q];
}

// Scan (all lanes; Hillis-Steele over kWG lanes in LDS with barriers between
// the steps supplied by the caller): tc[j] <- cnt[0] + ... + cnt[j].  A step
// with distance d first reads (ent_scan_read), then, after a barrier, adds.
CJD_HD inline void ent_scan_init(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J) return;
  for (int q = 0; q < 4; ++q) sh.tc[tid][q] = sh.u.a.cnt[tid][q];
}
CJD_HD inline void ent_scan_read(const EntShared& sh, int tid, int d, int32_t* v) {
  for (int q = 0; q < 4; ++q) v[q] = tid >= d ? sh.tc[tid - d][q] : 0;
}
CJD_HD inline void ent_scan_write(EntShared& sh, int tid, const int32_t* v) {
  for (int q = 0; q < 4; ++q) sh.tc[tid][q] += v[q];
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

// Aggregate of this item (tc scanned: relative cumulative counts).
CJD_HD inline void ent_aggregate(const EntShared& sh, const EntCtx& c, LookbackRec* me) {
  me->lead_p = sh.ep[0];
  me->lead_s = sh.es[0];
  me->agg_p = sh.tp[c.J - 1];
  me->agg_s = sh.ts[c.J - 1];
  for (int q = 0; q < 4; ++q) me->agg_c[q] = sh.tc[c.J - 1][q];
}

// Inclusive state of this item: the true exit of its last tile and the
// segment's cumulative counts there (tc absolute).
CJD_HD inline void ent_inclusive(const EntShared& sh, const EntCtx& c, LookbackRec* me) {
  me->incl_p = sh.tp[c.J - 1];
  me->incl_s = sh.ts[c.J - 1];
  for (int q = 0; q < 4; ++q) me->incl_c[q] = sh.tc[c.J - 1][q];
}

// After the lookback: whether the true entry differs from the assumed one
// (then tile 0's exit is set to it and the fix-up rounds and the scan run
// again), and the absolute counts (every lane adds the entry's counts).
CJD_HD inline bool ent_entry_differs(const EntShared& sh, const EntCtx& c) {
  return c.lead && (sh.entry_p != sh.ep[0] || sh.entry_s != sh.es[0]);
}
CJD_HD inline void ent_absolute(EntShared& sh, int tid, const EntCtx& c) {
  if (tid >= c.J || !c.lead) return;
  for (int q = 0; q < 4; ++q) sh.tc[tid][q] += sh.entry_c[q];
}

// Phase 5: final decode of local tile tid + lead from its true entry.
CJD_HD inline void ent_phase5(EntShared& sh, int tid, const EntCtx& c, int16_t* coefs, uint32_t* status) {
  const int fj = tid + c.lead;
  if (fj >= c.J) return;
  const int nh = c.nh;
  const uint32_t t = c.tfirst + (uint32_t)fj;
  uint32_t* st = status + c.im->error_index;
  uint32_t p;
  int slot;
  int32_t cp[4];
  if (fj == 0) {
    p = 0;
    slot = 0;
    for (int q = 0; q < 4; ++q) cp[q] = 0;
  } else {
    p = sh.tp[fj - 1];
    slot = (int)sh.ts[fj - 1];
    for (int q = 0; q < 4; ++q) cp[q] = sh.tc[fj - 1][q];
  }
  const uint32_t xp = sh.tp[fj], xs = sh.ts[fj];
  int32_t cx[4];
  for (int q = 0; q < 4; ++q) cx[q] = sh.tc[fj][q];
  const uint32_t end = c.tile_end(t);
  int64_t bidx = cp[0];
  int dc0 = cp[1], dc1 = cp[2], dc2 = cp[3];
  int16_t* blk = sh.u.coef[tid];
  int err = 0;
  bool stopped = false;
  BitReader br;
  br.seek(c.w, c.nwords, p);
  int coef = -1;
  for (bool done = false; !done;) {
    br.rotate();
    for (int u = 0; u < kPeriod; ++u) {
      if (coef < 0) {
        if (br.pos >= end) {
          done = true;
          break;
        }
        if (bidx >= (int64_t)c.sg.nblocks) {
          stopped = true;
          done = true;
          break;
        }
        __builtin_memset(blk, 0, 128);
      }
      const int s0 = slot, comp = c.comp(s0);
      int d, wp, wv;
      const bool ended = sym_step(br, sh, c.slotinfo, nh, slot, coef, d, wp, wv, err);
      if (err) {
        done = true;
        break;
      }
      dc0 += comp == 0 ? d : 0;
      dc1 += comp == 1 ? d : 0;
      dc2 += comp == 2 ? d : 0;
      if (wp >= 0) blk[wp] = (int16_t)wv;
      if (ended) {
        blk[0] = (int16_t)(comp == 0 ? dc0 : comp == 1 ? dc1 : dc2);
        __builtin_memcpy(coefs + (c.sg.first_block + bidx) * 64, blk, 128);
        ++bidx;
      }
    }
  }
  p = br.pos;
  if (err) {
    status_or(st, kStatusSyntax);
  } else if (!stopped) {
    if (p != xp || (uint32_t)slot != xs || bidx != (int64_t)cx[0] || dc0 != cx[1] || dc1 != cx[2] || dc2 != cx[3])
      status_or(st, kStatusInconsistent);
    if (t + 1 == c.sg.ntiles && bidx < (int64_t)c.sg.nblocks) status_or(st, kStatusTruncated);
  }
}

// Serial decode of one whole segment from its known start (reference order,
// src/decodeScanGPU.cu:435-451).
CJD_HD inline void ent_serial_segment(const uint32_t* words, const ImageDesc* images, const SegmentDesc& sg,
                                      const HuffTable* tables, int16_t* coefs, uint32_t* status) {
  const ImageDesc& im = images[sg.image];
  const HuffTable* tab = tables + im.table_base;
  BitReader br;
  br.seek(words + sg.word_base, sg.nwords, 0);
  int slot = 0, dc[3] = {0, 0, 0};
  int16_t blk[64];
  for (uint32_t b = 0; b < sg.nblocks; ++b) {
    for (int k = 0; k < 64; ++k) blk[k] = 0;
    int d, err = 0;
    decode_block<true>(br, tab[im.slot_dc[slot]], tab[im.slot_ac[slot]], d, err, blk);
    if (err) {
      status_or(status + im.error_index, kStatusSyntax);
      return;
    }
    const int comp = im.slot_comp[slot];
    dc[comp] += d;
    blk[0] = (int16_t)dc[comp];
    int16_t* g = coefs + (sg.first_block + b) * 64;
    for (int k = 0; k < 64; ++k) g[k] = blk[k];
    if (++slot == im.nslots) slot = 0;
  }
}

}  // namespace cjd
