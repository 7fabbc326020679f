// Progressive-scan decoder (CJD_EXT_PROGRESSIVE) of the MI355X JPEG decoder:
// __host__ __device__ functions run by k_prog / k_prog_prepare (prog.hip)
// and, lane by lane, by the CPU emulator of the tests.
//
// The reference decodes baseline scans only (src/format.cu:198-200 rejects
// SOF2); the procedures are T.81 G.1.2.  A progressive image's coefficients
// are built up over its scans in the dense coefficient buffer: k_prog_prepare
// zeroes them and gives every block a dense record (the pixel stage reads
// dense blocks from that buffer), then k_prog runs one lane per restart
// interval of a scan.  Scans are ordered into levels on the host: a scan's
// level is one more than that of the last earlier scan that codes a
// coefficient it codes (same component, overlapping band), so the scans of
// one level touch disjoint coefficients and run side by side, one launch per
// level.  A lane writes only the coefficients of its own band (16-bit
// stores), so lanes of one level never write the same bytes.
#pragma once
#include "entropy_core.h"

namespace cjd {

// MSB-first reader of one unstuffed restart interval: a 64-bit accumulator
// fed one 32-bit word at a time; words past the interval's buffer read as
// 0xFFFFFFFF (the 0xFF padding a decoder reads past the data,
// src/decodeScanGPU.cu:32).  pos counts the bits consumed.
struct ProgBits {
  const uint32_t* w;
  uint32_t nwords, wi, pos;
  uint64_t acc;
  int nacc;
  CJD_HD inline void init(const uint32_t* words, uint32_t n) {
    w = words;
    nwords = n;
    wi = 0;
    pos = 0;
    acc = 0;
    nacc = 0;
    fill();
  }
  CJD_HD inline void fill() {
    while (nacc <= 32) {
      const uint32_t x = wi < nwords ? bswap32(w[wi]) : 0xFFFFFFFFu;
      ++wi;
      acc |= (uint64_t)x << (32 - nacc);
      nacc += 32;
    }
  }
  CJD_HD inline uint32_t peek() const { return (uint32_t)(acc >> 32); }
  CJD_HD inline void skip(int n) {
    acc <<= n;
    nacc -= n;
    pos += (uint32_t)n;
    fill();
  }
  CJD_HD inline uint32_t get(int n) {
    if (!n) return 0;
    const uint32_t v = peek() >> (32 - n);
    skip(n);
    return v;
  }
};

// Coefficient-buffer index of block j of a task (j counts the blocks of its
// units in decode order) and, in ci, the index of its component in the scan.
CJD_HD inline int64_t prog_block(const ProgTask& t, const ImageDesc& im, uint32_t j, int& ci) {
  if (t.ncomp == 1) {
    const int64_t a = (int64_t)t.u0 + j, by = a / t.bw, bx = a - by * t.bw;
    ci = 0;
    return im.coef_base + ((by / t.v) * im.mcus_x + bx / t.h) * im.nslots + im.comp_first_slot[t.comps[0]] +
           (by % t.v) * t.h + bx % t.h;
  }
  const uint32_t u = t.u0 + j / t.bpu;
  uint32_t r = j % t.bpu;
  int c = t.comps[0];
  for (ci = 0; ci < t.ncomp; ++ci) {
    c = t.comps[ci];
    const uint32_t n = (uint32_t)im.comp_sx[c] * im.comp_sy[c];
    if (r < n) break;
    r -= n;
  }
  return im.coef_base + (int64_t)u * im.nslots + im.comp_first_slot[c] + r;
}

// zigzag position of each natural index (the inverse of kDeZigZag)
constexpr uint8_t kZigZagOf[64] = {0,  1,  5,  6,  14, 15, 27, 28, 2,  4,  7,  13, 16, 26, 29, 42,
                                   3,  8,  12, 17, 25, 30, 41, 43, 9,  11, 18, 24, 31, 40, 44, 53,
                                   10, 19, 23, 32, 39, 45, 52, 54, 20, 22, 33, 38, 46, 51, 55, 60,
                                   21, 34, 37, 47, 50, 56, 59, 61, 35, 36, 48, 49, 57, 58, 62, 63};

// The 128 bytes of one coefficient block as 32 words (16-byte loads on the
// device; the block is 128-byte aligned).
CJD_HD inline void prog_load_block(const int16_t* q, uint32_t (&w)[32]) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4* p = reinterpret_cast<const uint4*>(q);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint4 v = p[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
#else
  memcpy(w, q, 128);
#endif
}

// Add delta to coefficient z (natural index) of block q whose value x has the
// sign of delta and no carry out of its 16 bits in the direction of delta
// (a refinement moves a magnitude away from zero): on the device an atomic
// add to the word that holds it (a lane of the same level may store the
// word's other coefficient at the same time), so the value need not be read.
// For a coefficient in the low half a negative delta borrows nothing from
// the high half, as x's own 16 bits are at least 0x8000 >= |delta|.
CJD_HD inline void prog_coef_add(int16_t* q, int z, int delta) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t* w = reinterpret_cast<uint32_t*>(q) + (z >> 1);
  __hip_atomic_fetch_add(w, (z & 1) ? (uint32_t)delta << 16 : (uint32_t)delta, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
#else
  q[z] = (int16_t)(q[z] + delta);
#endif
}

// q[0] |= bit, as prog_coef_add does it (the DC coefficient is the low half).
CJD_HD inline void prog_dc_or(int16_t* q, int bit) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(q), (uint32_t)bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
  q[0] = (int16_t)(q[0] | bit);
#endif
}

// Refinement state of one block over the band: bit z of nz / neg / has is set
// where the coefficient at zigzag position z is nonzero / negative / already
// has bit Al, from the block as it was when the lane reached it (a scan
// changes each coefficient at most once, after passing it).
struct ProgMasks {
  uint64_t nz, neg, has;
};
CJD_HD inline ProgMasks prog_masks(const uint32_t (&w)[32], uint64_t band, int p1) {
  ProgMasks m{0, 0, 0};
#pragma unroll
  for (int n = 0; n < 64; ++n) {
    const int16_t x = (int16_t)(w[n >> 1] >> ((n & 1) * 16));
    const uint64_t bit = 1ull << kZigZagOf[n];
    m.nz |= x ? bit : 0;
    m.neg |= x < 0 ? bit : 0;
    m.has |= (x & p1) ? bit : 0;
  }
  m.nz &= band;
  return m;
}

// Correction bits of the nonzero coefficients in `span` (a set of zigzag
// positions), read in position order: a 1 moves the coefficient 2^Al away
// from zero unless it has that bit already.
CJD_HD inline void prog_corrections(ProgBits& br, const ProgMasks& m, uint64_t span, int16_t* q, int p1) {
  uint64_t todo = m.nz & span;
  while (todo) {
    const int n = __builtin_popcountll(todo) < 32 ? __builtin_popcountll(todo) : 32;
    uint32_t bits = br.get(n) << (32 - n);  // MSB first
    for (int i = 0; i < n; ++i, bits <<= 1) {
      const int z = __builtin_ctzll(todo);
      todo &= todo - 1;
      if ((bits & 0x80000000u) && !((m.has >> z) & 1)) prog_coef_add(q, kDeZigZag[z], ((m.neg >> z) & 1) ? -p1 : p1);
    }
  }
}

// Zigzag positions [a, b] as a mask (0 <= a, b <= 63; empty when a > b).
CJD_HD inline uint64_t prog_span(int a, int b) {
  if (a > b) return 0;
  const uint64_t hi = b >= 63 ? ~0ull : (1ull << (b + 1)) - 1;
  return hi & (~0ull << a);
}

// Decode one task (a restart interval of a scan) into coefs.  DC first: the
// DC difference as in baseline, the predictor stored << Al; DC refine: one
// raw bit per block ORed in at bit Al; AC first: run/size symbols over
// Ss..Se with values << Al, EOBn (size 0, run r < 15) ending this block and
// the next 2^r + (r extra bits) - 1 blocks; AC refine: a size-1 symbol's
// sign bit makes a new coefficient +-2^Al at the (run + 1)-th coefficient of
// the band still zero, every nonzero coefficient passed on the way takes a
// correction bit, ZRL passes 16 zero coefficients, and EOBn ends the symbols
// of this block and the next blocks of the run, whose nonzero coefficients
// still take correction bits.  An invalid code, a refinement size other than
// 0 or 1, a run past Se, or (for an interval followed by an RSTn) data that
// does not end in the interval's last byte sets kStatusSyntax.
CJD_HD inline void prog_task(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                             const ProgTask& t, const HuffTable* tables, int16_t* coefs, uint32_t* status,
                             uint32_t epoch) {
  const SegmentDesc& sg = segs[t.seg];
  const ImageDesc& im = images[sg.image];
  const HuffTable* tab = tables + (size_t)sg.tset * 4;
  ProgBits br;
  br.init(words + sg.word_base, sg.nwords);
  const uint32_t nb = t.nunits * t.bpu;
  const int p1 = 1 << t.al;
  bool err = false;
  if (t.ss == 0 && t.ah == 0) {  // DC first
    uint32_t pred[3] = {0, 0, 0};
    for (uint32_t j = 0; j < nb && !err; ++j) {
      int ci;
      const int64_t g = prog_block(t, im, j, ci);
      const uint32_t bits = br.peek();
      int len;
      const int sym = huff_sym(tab[t.dcs[ci]], bits, len);
      if (!len) {
        err = true;
        break;
      }
      const int size = sym & 15;
      pred[ci] += (uint32_t)extend_value(bits, len, size);
      br.skip(len + size);
      coefs[g * 64] = (int16_t)(uint16_t)(pred[ci] << t.al);
    }
  } else if (t.ss == 0) {  // DC refine
    for (uint32_t j = 0; j < nb; ++j) {
      int ci;
      const int64_t g = prog_block(t, im, j, ci);
      if (br.get(1)) prog_dc_or(coefs + g * 64, p1);
    }
  } else if (t.ah == 0) {  // AC first
    const HuffTable& act = tab[t.ac];
    uint32_t eobrun = 0;
    for (uint32_t j = 0; j < nb && !err; ++j) {
      int ci;
      const int64_t g = prog_block(t, im, j, ci);
      if (eobrun) {
        --eobrun;
        continue;
      }
      int16_t* q = coefs + g * 64;
      for (int k = t.ss; k <= t.se;) {
        const uint32_t bits = br.peek();
        int len;
        const int sym = huff_sym(act, bits, len);
        if (!len) {
          err = true;
          break;
        }
        const int run = sym >> 4, size = sym & 15;
        if (size) {
          k += run;
          if (k > t.se) {
            err = true;
            break;
          }
          q[kDeZigZag[k++]] = (int16_t)(uint16_t)((uint32_t)extend_value(bits, len, size) << t.al);
          br.skip(len + size);
        } else if (run == 15) {
          br.skip(len);
          k += 16;
          if (k > t.se + 1) {
            err = true;
            break;
          }
        } else {
          br.skip(len);
          eobrun = (1u << run) - 1u + br.get(run);
          break;
        }
      }
    }
  } else {  // AC refine
    const HuffTable& act = tab[t.ac];
    const uint64_t band = prog_span(t.ss, t.se);
    uint32_t eobrun = 0;
    uint32_t nxt[32];
    int ci;
    if (nb) prog_load_block(coefs + prog_block(t, im, 0, ci) * 64, nxt);
    for (uint32_t j = 0; j < nb && !err; ++j) {
      const int64_t g = prog_block(t, im, j, ci);
      int16_t* q = coefs + g * 64;
      uint32_t cur[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) cur[i] = nxt[i];
      if (j + 1 < nb) prog_load_block(coefs + prog_block(t, im, j + 1, ci) * 64, nxt);  // in flight meanwhile
      const ProgMasks m = prog_masks(cur, band, p1);
      int k = t.ss;
      if (!eobrun) {
        while (k <= t.se) {
          const uint32_t bits = br.peek();
          int len;
          const int sym = huff_sym(act, bits, len);
          if (!len) {
            err = true;
            break;
          }
          const int run = sym >> 4, size = sym & 15;
          if (size > 1) {
            err = true;
            break;
          }
          // a size-1 symbol's extra bit is the new coefficient's sign
          const int v = extend_value(bits, len, size);
          br.skip(len + size);
          if (!size && run != 15) {
            eobrun = (1u << run) + br.get(run);
            break;
          }
          // the target: the (run + 1)-th coefficient from k on that is zero;
          // the nonzero ones before it take correction bits
          uint64_t zeros = ~m.nz & prog_span(k, t.se);
          for (int r = 0; r < run && zeros; ++r) zeros &= zeros - 1;
          if (!zeros) {  // the band ends first
            prog_corrections(br, m, prog_span(k, t.se), q, p1);
            k = t.se + 1;
            err = size != 0;
            break;
          }
          const int tz = __builtin_ctzll(zeros);
          prog_corrections(br, m, prog_span(k, tz - 1), q, p1);
          if (size) q[kDeZigZag[tz]] = (int16_t)(v * p1);
          k = tz + 1;
        }
        if (err) break;
      }
      if (eobrun) {
        prog_corrections(br, m, prog_span(k, t.se), q, p1);
        --eobrun;
      }
    }
  }
  // an interval followed by an RSTn must end in its last data byte (the
  // serial decoders byte-align there and read the marker)
  if (!err && !t.last && (br.pos + 7) / 8 != sg.nbits / 8) err = true;
  if (err) status_or(status + im.error_index, kStatusSyntax, epoch);
}

// k_prog_prepare for block i of the progressive images (n of them): its
// coefficients zeroed and a dense record.
CJD_HD inline void prog_prepare_block(const ProgImage* pimgs, uint32_t n, int64_t i, int16_t* coefs,
                                      BlockRec* recs) {
  uint32_t lo = 0, hi = n;  // the last image with block0 <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (pimgs[mid].block0 <= i) lo = mid;
    else hi = mid;
  }
  const int64_t g = pimgs[lo].coef_base + (i - pimgs[lo].block0);
  int16_t* q = coefs + g * 64;
  for (int k = 0; k < 64; ++k) q[k] = 0;
  BlockRec r;
  r.eidx = 0;
  r.nw = kRecDense;
  r.dc = 0;
  recs[g] = r;
}

}  // namespace cjd
