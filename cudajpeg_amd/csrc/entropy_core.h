// Entropy-decode core of the MI355X baseline-JPEG decoder: the symbol decoder
// and the phases of the tile kernel, written as __host__ __device__ functions
// so that the same code runs in k_entropy (entropy.hip) and, lane by lane, in
// the CPU emulator used by the tests.
//
// Algorithm (replaces src/entropyRLEdecodeGPU.cu:7-238 and the host serial
// decode of src/decodeScanGPU.cu:435-451):
//  1. Every tile of S bits is decoded speculatively by one lane per MCU slot h
//     ("hypothesis" h: a block of slot h starts at the tile's first bit).  Each
//     lane remembers its first kListLen block starts (bit, slot, DC prefix) and
//     its exit: the first block start at or after the tile end, with the block
//     count and DC sums on the way.
//  2. Link: from every hypothesis exit of tile t a lane decodes into tile t+1
//     until it reaches a remembered block start of a hypothesis h' of t+1
//     (the trajectories have synchronised): sigma_t(h) = h' plus the blocks /
//     DC sums added in tile t+1.
//  3. These maps compose associatively: pointer jumping in LDS gives, for
//     every tile of the workgroup, the state reached from each hypothesis of
//     the lead-in tile; a decoupled lookback across workgroups supplies the
//     true state of the lead-in tile.
//  4. Every tile is decoded once more from its true entry (bit, slot, block
//     index, DC predictors), writing quantised coefficients.
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
// a 64-bit accumulator fed one 32-bit word at a time from a 4-word chunk; the
// following chunk is already in flight (a 16-byte load issued one chunk ahead),
// so a lane decoding symbols does not wait on memory.  Words past the
// segment read as 0xFFFFFFFF (the reference pads with 0xFF beyond the end,
// src/decodeScanGPU.cu:32).
struct BitReader {
  const uint32_t* w;
  uint32_t nwords;
  uint64_t acc;   // next bits, MSB-aligned; nacc of them valid (> 32 after refill)
  int nacc;
  uint32_t q0, q1, q2, q3;  // current chunk (byte-swapped)
  uint32_t n0, n1, n2, n3;  // next chunk
  uint32_t qi;    // next word of the current chunk to append
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
    a = bswap32(v.x);
    b = bswap32(v.y);
    c = bswap32(v.z);
    d = bswap32(v.w);
#else
    a = bswap32(w[wi]);
    b = bswap32(w[wi + 1]);
    c = bswap32(w[wi + 2]);
    d = bswap32(w[wi + 3]);
#endif
  }
  CJD_HD inline void refill() {
    while (nacc <= 32) {
      const uint32_t word = qi == 0 ? q0 : qi == 1 ? q1 : qi == 2 ? q2 : q3;
      acc |= (uint64_t)word << (32 - nacc);
      nacc += 32;
      if (++qi == 4) {
        q0 = n0;
        q1 = n1;
        q2 = n2;
        q3 = n3;
        chunk += 4;
        qi = 0;
        load(chunk + 4, n0, n1, n2, n3);
      }
    }
  }
  CJD_HD inline void seek(const uint32_t* words, uint32_t n, uint32_t p) {
    w = words;
    nwords = n;
    const uint32_t wi = p >> 5;
    chunk = wi & ~3u;
    load(chunk, q0, q1, q2, q3);
    load(chunk + 4, n0, n1, n2, n3);
    qi = wi - chunk;
    acc = 0;
    nacc = 0;
    pos = p;
    refill();
    const int o = (int)(p & 31);
    acc <<= o;
    nacc -= o;
    refill();
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
struct EntShared {
  HuffTable tab[4];
  uint8_t dz[64];                     // zigzag index -> natural index
  uint32_t exit_p[kWG];
  uint32_t exit_s[kWG];
  int32_t tot[kWG][4];   // blocks, dc0, dc1, dc2 of each hypothesis
  uint32_t nlist[kWG];
  union {
    struct {
      uint32_t p[kWG][kListLen];
      uint32_t s_dc0[kWG][kListLen];  // slot | dc0 << 16
      uint32_t dc12[kWG][kListLen];   // dc1 | dc2 << 16
    } l;
    int16_t coef[kWG][64];            // final pass: one block per lane
  } u;
  int32_t lk_h[kWG];                  // link of (tile j, hypothesis h) into tile j+1: hypothesis or -1
  int32_t lk_c[kWG][4];               //   ... blocks, dc0, dc1, dc2 gained in tile j+1
  uint32_t lk_p[kWG], lk_s[kWG];      //   ... exit of tile j+1 reached when the link is -1
  int32_t mh[2][kWG];                 // prefix compositions (pointer jumping)
  int32_t mc[2][kWG][4];
  int32_t ex_h[kWG];                  // true state at the exit of each local tile:
  uint32_t ex_p[kWG], ex_s[kWG];      //   hypothesis (-1: none, -2: not yet known), bit, slot,
  int32_t ex_c[kWG][4];               //   cumulative blocks / DC sums of the segment
  int32_t entry_h;                    // true state at the exit of local tile 0
  uint32_t entry_p, entry_s;
  int32_t entry_c[4];
  uint32_t item;
  int32_t buf;                        // which mh/mc buffer holds the prefixes
  int32_t need_walk;
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
  CJD_HD const HuffTable& dct(const EntShared& sh, int slot) const {
    return sh.tab[(slotinfo >> (4 * slot)) & 1];
  }
  CJD_HD const HuffTable& act(const EntShared& sh, int slot) const {
    return sh.tab[2 + ((slotinfo >> (4 * slot + 1)) & 1)];
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

// Phase 1: speculative decode of (local tile j, hypothesis h) = lane tid.
CJD_HD inline void ent_phase1(EntShared& sh, int tid, const EntCtx& c) {
  const int j = tid / c.nh, h = tid - j * c.nh;
  if (j >= c.J) return;
  const uint32_t t = c.tfirst + (uint32_t)j;
  BitReader br;
  br.seek(c.w, c.nwords, t * c.S);
  int slot = (t == 0) ? 0 : h, coef = -1, err = 0;
  const uint32_t end = c.tile_end(t);
  int32_t nb = 0, dc[3] = {0, 0, 0};
  for (;;) {
    if (coef < 0) {
      const uint32_t p = br.pos;
      if (p >= end) break;
      if (nb < kListLen) {
        sh.u.l.p[tid][nb] = p;
        sh.u.l.s_dc0[tid][nb] = (uint32_t)slot | ((uint32_t)(uint16_t)(int16_t)dc[0] << 16);
        sh.u.l.dc12[tid][nb] = (uint32_t)(uint16_t)(int16_t)dc[1] | ((uint32_t)(uint16_t)(int16_t)dc[2] << 16);
      }
      ++nb;
    }
    const int s0 = slot;
    int d, wp, wv;
    sym_step(br, sh, c.slotinfo, c.nh, slot, coef, d, wp, wv, err);
    add_dc(dc, c.comp(s0), d);
  }
  sh.exit_p[tid] = br.pos;
  sh.exit_s[tid] = (uint32_t)slot;
  sh.tot[tid][0] = nb;
  sh.tot[tid][1] = dc[0];
  sh.tot[tid][2] = dc[1];
  sh.tot[tid][3] = dc[2];
  sh.nlist[tid] = nb < kListLen ? (uint32_t)nb : (uint32_t)kListLen;
}

// Decode from state (p, slot) into local tile j until the trajectory lands on
// a remembered block start of one of tile j's hypotheses (hit, list index) or
// reaches the tile end.  On a hit, cnt gets the blocks / DC sums of the whole
// tile j along the path and (p, slot) its exit; otherwise the tile was decoded
// directly and (p, slot) is the concrete exit.  Returns the hypothesis hit
// (-1 for none).
CJD_HD inline int ent_resync(const EntShared& sh, const EntCtx& c, int j, uint32_t& p, int& slot, int32_t* cnt) {
  const int nh = c.nh;
  const uint32_t end = c.tile_end(c.tfirst + (uint32_t)j);
  int32_t rb = 0, r[3] = {0, 0, 0};
  uint64_t cur = 0;  // 4-bit list cursor per hypothesis (kept in a register)
  int hit = -1, hit_i = 0;
  const int base = j * nh;
  BitReader br;
  br.seek(c.w, c.nwords, p);
  int coef = -1, err = 0;
  for (;;) {
    if (coef >= 0) {  // inside a block: one more symbol
      const int s0 = slot;
      int d, wp, wv;
      sym_step(br, sh, c.slotinfo, nh, slot, coef, d, wp, wv, err);
      add_dc(r, c.comp(s0), d);
      continue;
    }
    p = br.pos;
    for (int k = 0; k < nh && hit < 0; ++k) {
      const int L = base + k;
      const int n = (int)sh.nlist[L];
      int q = (int)((cur >> (4 * k)) & 15);
      while (q < n && sh.u.l.p[L][q] < p) ++q;
      cur = (cur & ~(15ull << (4 * k))) | ((uint64_t)q << (4 * k));
      if (q < n && sh.u.l.p[L][q] == p && (int)(sh.u.l.s_dc0[L][q] & 0xFF) == slot) {
        hit = k;
        hit_i = q;
      }
    }
    if (hit >= 0 || p >= end) break;
    ++rb;
    const int s0 = slot;
    int d, wp, wv;
    sym_step(br, sh, c.slotinfo, nh, slot, coef, d, wp, wv, err);
    add_dc(r, c.comp(s0), d);
  }
  cnt[0] = rb;
  cnt[1] = r[0];
  cnt[2] = r[1];
  cnt[3] = r[2];
  if (hit >= 0) {
    const int L = base + hit;
    const uint32_t sd = sh.u.l.s_dc0[L][hit_i], d12 = sh.u.l.dc12[L][hit_i];
    cnt[0] += sh.tot[L][0] - hit_i;
    cnt[1] += sh.tot[L][1] - (int32_t)(int16_t)(sd >> 16);
    cnt[2] += sh.tot[L][2] - (int32_t)(int16_t)(d12 & 0xFFFF);
    cnt[3] += sh.tot[L][3] - (int32_t)(int16_t)(d12 >> 16);
    p = sh.exit_p[L];
    slot = (int)sh.exit_s[L];
    return hit;
  }
  // landed at or past the end of tile j: it may coincide with an exit
  for (int k = 0; k < nh; ++k) {
    const int L = base + k;
    if (sh.exit_p[L] == p && (int)sh.exit_s[L] == slot) return k;
  }
  return -1;
}

// Phase 2: link the exit of (j, h) = lane tid into tile j+1.
CJD_HD inline void ent_phase2(EntShared& sh, int tid, const EntCtx& c) {
  const int j = tid / c.nh;
  if (j + 1 >= c.J) return;
  uint32_t p = sh.exit_p[tid];
  int slot = (int)sh.exit_s[tid];
  int32_t cnt[4];
  const int hit = ent_resync(sh, c, j + 1, p, slot, cnt);
  sh.lk_h[tid] = hit;
  sh.lk_p[tid] = p;
  sh.lk_s[tid] = (uint32_t)slot;
  for (int q = 0; q < 4; ++q) sh.lk_c[tid][q] = cnt[q];
}

// Phase 3 init: F_j = link from tile j-1 (identity for j = 0) into buffer 1.
CJD_HD inline void ent_phase3_init(EntShared& sh, int tid, const EntCtx& c) {
  const int j = tid / c.nh, h = tid - j * c.nh;
  if (j >= c.J) return;
  if (j == 0) {
    sh.mh[1][tid] = h;
    for (int q = 0; q < 4; ++q) sh.mc[1][tid][q] = 0;
  } else {
    const int L = (j - 1) * c.nh + h;
    sh.mh[1][tid] = sh.lk_h[L];
    for (int q = 0; q < 4; ++q) sh.mc[1][tid][q] = sh.lk_c[L][q];
  }
}

// Phase 3 step (pointer jumping): G_j <- G_j o G_{j-d}, from buffer `buf`
// into buffer buf^1.  G_j(h) = state in local tile j reached from hypothesis h
// of local tile 0.
CJD_HD inline void ent_phase3_step(EntShared& sh, int tid, const EntCtx& c, int d, int buf) {
  const int j = tid / c.nh, h = tid - j * c.nh;
  if (j >= c.J) return;
  const int nb = buf ^ 1;
  if (j >= d) {
    const int L0 = (j - d) * c.nh + h;
    const int h1 = sh.mh[buf][L0];
    if (h1 < 0) {
      sh.mh[nb][tid] = -1;
      for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = 0;
    } else {
      const int L1 = j * c.nh + h1;
      sh.mh[nb][tid] = sh.mh[buf][L1];
      for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = sh.mc[buf][L0][q] + sh.mc[buf][L1][q];
    }
  } else {
    sh.mh[nb][tid] = sh.mh[buf][tid];
    for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = sh.mc[buf][tid][q];
  }
}

// Phase 4a: publish this item's aggregate (the map from local tile 0's
// hypotheses to the last tile).  For the first item of a segment the state at
// tile 0 is known (hypothesis 0 decodes from the segment start): it is set and
// true is returned.
CJD_HD inline bool ent_publish_aggregate(EntShared& sh, const EntCtx& c, LookbackRec* me) {
  const int Lq = (c.J - 1) * c.nh;
  const int buf = sh.buf;
  if (!c.lead) {
    sh.entry_h = 0;
    sh.entry_p = sh.exit_p[0];
    sh.entry_s = sh.exit_s[0];
    for (int q = 0; q < 4; ++q) sh.entry_c[q] = sh.tot[0][q];
    return true;
  }
  for (int k = 0; k < c.nh; ++k) {
    me->agg_h[k] = (int8_t)sh.mh[buf][Lq + k];
    for (int q = 0; q < 4; ++q) me->agg_c[k][q] = sh.mc[buf][Lq + k][q];
  }
  return false;
}

// Composition accumulator of the lookback: acc = agg_{w-1} o ... o agg_q.
struct LookAcc {
  int h[kMaxSlots];
  int32_t c[kMaxSlots][4];
  bool ident;
  CJD_HD void identity(int nh) {
    ident = true;
    for (int k = 0; k < nh; ++k) {
      h[k] = k;
      for (int q = 0; q < 4; ++q) c[k][q] = 0;
    }
  }
  // acc <- acc o agg(r)
  CJD_HD void absorb(const LookbackRec& r, int nh) {
    ident = false;
    int nh_[kMaxSlots];
    int32_t nc[kMaxSlots][4];
    for (int k = 0; k < nh; ++k) {
      const int h1 = r.agg_h[k];
      if (h1 < 0 || h[h1] < 0) {
        nh_[k] = -1;
        for (int q = 0; q < 4; ++q) nc[k][q] = 0;
      } else {
        nh_[k] = h[h1];
        for (int q = 0; q < 4; ++q) nc[k][q] = r.agg_c[k][q] + c[h1][q];
      }
    }
    for (int k = 0; k < nh; ++k) {
      h[k] = nh_[k];
      for (int q = 0; q < 4; ++q) c[k][q] = nc[k][q];
    }
  }
  // Entry state (at the exit of local tile 0) from a predecessor's inclusive
  // state; false when it cannot be derived through the composed aggregates.
  CJD_HD bool apply(const LookbackRec& r, const EntShared& sh, int& eh, uint32_t& ep, uint32_t& es,
                    int32_t* ec) const {
    if (ident) {
      eh = r.incl_h;
      ep = r.incl_p;
      es = r.incl_s;
      for (int q = 0; q < 4; ++q) ec[q] = r.incl_c[q];
      return true;
    }
    const int hb = r.incl_h;
    if (hb < 0 || h[hb] < 0) return false;
    eh = h[hb];
    ep = sh.exit_p[eh];  // local tile 0, hypothesis eh
    es = sh.exit_s[eh];
    for (int q = 0; q < 4; ++q) ec[q] = r.incl_c[q] + c[hb][q];
    return true;
  }
};

// Phase 4b (all lanes): true exit state of local tile j from the prefix
// compositions; -2 where a link on the path did not resolve.
CJD_HD inline void ent_states(EntShared& sh, int tid, const EntCtx& c) {
  const int j = tid;
  if (j >= c.J) return;
  const int eh = sh.entry_h;
  if (j == 0) {
    sh.ex_h[0] = eh;
    sh.ex_p[0] = sh.entry_p;
    sh.ex_s[0] = sh.entry_s;
    for (int q = 0; q < 4; ++q) sh.ex_c[0][q] = sh.entry_c[q];
    return;
  }
  const int buf = sh.buf;
  if (eh >= 0) {
    const int L = j * c.nh + eh;
    const int hj = sh.mh[buf][L];
    if (hj >= 0) {
      sh.ex_h[j] = hj;
      sh.ex_p[j] = sh.exit_p[j * c.nh + hj];
      sh.ex_s[j] = sh.exit_s[j * c.nh + hj];
      for (int q = 0; q < 4; ++q) sh.ex_c[j][q] = sh.entry_c[q] + sh.mc[buf][L][q];
      return;
    }
  }
  sh.ex_h[j] = -2;
  sh.need_walk = 1;
}

// Phase 4c (one lane): fill the states the compositions could not give by
// walking the tiles in order: a link where one exists, else the concrete exit
// recorded by the link, else a resync decode from a concrete state.  Once a
// composition fails every later one fails too (-1 propagates), so the walk
// runs from the first unknown tile to the last.
CJD_HD inline void ent_walk(EntShared& sh, const EntCtx& c) {
  for (int j = 1; j < c.J; ++j) {
    if (sh.ex_h[j] != -2) continue;
    const int hp = sh.ex_h[j - 1];
    int32_t cnt[4];
    if (hp >= 0) {
      const int L = (j - 1) * c.nh + hp;
      sh.ex_h[j] = sh.lk_h[L];
      sh.ex_p[j] = sh.lk_p[L];
      sh.ex_s[j] = sh.lk_s[L];
      for (int q = 0; q < 4; ++q) cnt[q] = sh.lk_c[L][q];
    } else {
      uint32_t p = sh.ex_p[j - 1];
      int slot = (int)sh.ex_s[j - 1];
      sh.ex_h[j] = ent_resync(sh, c, j, p, slot, cnt);
      sh.ex_p[j] = p;
      sh.ex_s[j] = (uint32_t)slot;
    }
    for (int q = 0; q < 4; ++q) sh.ex_c[j][q] = sh.ex_c[j - 1][q] + cnt[q];
  }
}

// Phase 4d: the inclusive state of this item (exit of its last tile).
CJD_HD inline void ent_inclusive(const EntShared& sh, const EntCtx& c, LookbackRec* me) {
  const int j = c.J - 1;
  me->incl_h = sh.ex_h[j];
  me->incl_p = sh.ex_p[j];
  me->incl_s = sh.ex_s[j];
  for (int q = 0; q < 4; ++q) me->incl_c[q] = sh.ex_c[j][q];
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
  if (!c.lead && fj == 0) {
    p = 0;
    slot = 0;
    for (int q = 0; q < 4; ++q) cp[q] = 0;
  } else {
    p = sh.ex_p[fj - 1];
    slot = (int)sh.ex_s[fj - 1];
    for (int q = 0; q < 4; ++q) cp[q] = sh.ex_c[fj - 1][q];
  }
  const int32_t* cx = sh.ex_c[fj];
  const uint32_t end = c.tile_end(t);
  int64_t bidx = cp[0];
  int dc0 = cp[1], dc1 = cp[2], dc2 = cp[3];
  int16_t* blk = sh.u.coef[tid];
  int err = 0;
  bool stopped = false;
  BitReader br;
  br.seek(c.w, c.nwords, p);
  int coef = -1;
  for (;;) {
    if (coef < 0) {
      if (br.pos >= end) break;
      if (bidx >= (int64_t)c.sg.nblocks) {
        stopped = true;
        break;
      }
      __builtin_memset(blk, 0, 128);
    }
    const int s0 = slot, comp = c.comp(s0);
    int d, wp, wv;
    const bool ended = sym_step(br, sh, c.slotinfo, nh, slot, coef, d, wp, wv, err);
    if (err) break;
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
  p = br.pos;
  if (err) {
    status_or(st, kStatusSyntax);
  } else if (!stopped) {
    if (p != sh.ex_p[fj] || (uint32_t)slot != sh.ex_s[fj] || bidx != (int64_t)cx[0] || dc0 != cx[1] ||
        dc1 != cx[2] || dc2 != cx[3])
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
