// Progressive-scan decoder (CJD_EXT_PROGRESSIVE) of the MI355X JPEG decoder:
// __host__ __device__ functions run by k_prog / k_prog_prepare (prog.hip)
// and by the CPU emulator of the tests.
//
// The reference decodes baseline scans only (src/format.cu:198-200 rejects
// SOF2); the procedures are T.81 G.1.2.  A progressive image's coefficients
// are built up over its scans in the dense coefficient buffer: k_prog_prepare
// zeroes them and gives every block a dense record (the pixel stage reads
// dense blocks from that buffer), then k_prog decodes each restart interval
// of a scan (a task) with one wave64.  Scans are ordered into levels on the
// host: a scan's level is one more than that of the last earlier scan that
// codes a coefficient it codes (same component, overlapping band), so the
// scans of one level touch disjoint coefficients and run side by side, one
// launch per level.  A task writes only the coefficients of its own band, as
// 16-bit stores (the DC refinement as an atomic OR on the DC's word), so
// tasks of one level never write the same bytes.
//
// The wave decodes in lockstep: the Huffman decode is the same in every lane,
// and lane z owns the coefficient at zigzag position z of the current block
// (its nonzero / sign / bit-Al masks are three ballots; the coefficient's new
// value is kept in the lane and stored once per block).  Memory waits are
// what a lone wave cannot hide, and on CDNA the vector-memory counter also
// counts stores, so no vector-memory wait is left in the per-block loop: the
// stream is read from a ring in LDS refilled 4 KB at a time, a refinement
// scan's blocks are fetched 64 at a time, one batch ahead, into registers and
// then LDS, and the Huffman lookups (wave-uniform addresses) are scalar
// loads, which wait on a counter of their own.  On the host (the emulator) one thread plays all 64
// lanes: the per-lane values are arrays and the lane loops run in full.
#pragma once
#include "entropy_core.h"

namespace cjd {

constexpr int kProgRingWords = 2048;  // stream ring (8 KB); refilled 1024 words at a time
constexpr int kProgRefill = 1024;
constexpr int kProgBlockWords = 80;   // words kept ahead of a block (AC first takes at most 63 x 31 + 30 bits)
constexpr int kProgBatch = 64;        // refinement blocks fetched at a time

// LDS of one k_prog workgroup (a local variable of the emulator's).
struct ProgShared {
  uint32_t ring[kProgRingWords];   // stream words as stored (big-endian bytes)
  alignas(16) int16_t blk[kProgBatch][64];  // a batch of refinement blocks, natural order
};

// Per-lane values: the lane's own on the device, all 64 lanes' on the host.
template <class T>
struct Lanes {
#if defined(__HIP_DEVICE_COMPILE__)
  T v;
  CJD_HD inline T& at(int) { return v; }
  CJD_HD inline const T& at(int) const { return v; }
#else
  T v[64];
  CJD_HD inline T& at(int l) { return v[l]; }
  CJD_HD inline const T& at(int l) const { return v[l]; }
#endif
};
#if defined(__HIP_DEVICE_COMPILE__)
#define PROG_LANES(l) for (int l = (int)__lane_id(), l##_once = 1; l##_once; l##_once = 0)
#define PROG_SYNC() __syncthreads()
#else
#define PROG_LANES(l) for (int l = 0; l < 64; ++l)
#define PROG_SYNC() ((void)0)
#endif
template <class F>
CJD_HD inline uint64_t prog_ballot(F&& pred) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(pred((int)__lane_id()));
#else
  uint64_t m = 0;
  for (int l = 0; l < 64; ++l) m |= pred(l) ? 1ull << l : 0;
  return m;
#endif
}

// MSB-first reader of one unstuffed restart interval from the LDS ring: a
// 64-bit accumulator fed one word at a time.  The ring holds words
// [loaded - kProgRingWords, loaded) of the interval; refill() (all lanes, at
// a block boundary) loads the next kProgRefill words when fewer than
// kProgBlockWords are left ahead of the reader.  Words past the interval's
// nwords read as 0xFFFFFFFF (the 0xFF padding a decoder reads past the data,
// src/decodeScanGPU.cu:32).  pos counts the bits consumed.
struct ProgBits {
  const uint32_t* w;
  uint32_t* ring;
  uint32_t nwords, wi, loaded, pos;
  uint64_t acc;
  int nacc;
  CJD_HD inline void refill() {
    if (wi + kProgBlockWords <= loaded) return;
    PROG_SYNC();
    PROG_LANES(l) {
      for (int i = 0; i < kProgRefill / 64; i += 4) {
        const uint32_t at = loaded + (uint32_t)(i * 64 + l * 4);
        uint32_t* d = ring + (at & (kProgRingWords - 1));
        if (at + 4 <= nwords) {
#if defined(__HIP_DEVICE_COMPILE__)
          *reinterpret_cast<uint4*>(d) = *reinterpret_cast<const uint4*>(w + at);
#else
          memcpy(d, w + at, 16);
#endif
        } else {
          d[0] = d[1] = d[2] = d[3] = 0xFFFFFFFFu;
        }
      }
    }
    PROG_SYNC();
    loaded += kProgRefill;
  }
  CJD_HD inline void init(const uint32_t* words, uint32_t n, uint32_t* r) {
    w = words;
    ring = r;
    nwords = n;
    wi = loaded = pos = 0;
    acc = 0;
    nacc = 0;
    refill();
    fill();
  }
  CJD_HD inline void fill() {
    while (nacc <= 32) {
      acc |= (uint64_t)bswap32(ring[wi & (kProgRingWords - 1)]) << (32 - nacc);
      ++wi;
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

// The blocks of a task in decode order, without divisions per block: a
// one-component scan walks the component's grid (bx, by) and maps a block to
// its MCU and slot with shifts (sampling factors are powers of two); a scan
// of several components walks MCUs, taking each MCU's slots of its
// components from the task's packed list (ProgTask::slots, ::cis).
struct ProgIter {
  int64_t base;
  uint32_t nslots, mcus_x;
  bool one;
  uint32_t bx, by, bw, hs, vs, slot0;  // one component
  uint32_t u, r, bpu, cis;             // several
  uint64_t slots;
  CJD_HD inline void init(const ProgTask& t, const ImageDesc& im) {
    base = im.coef_base;
    nslots = (uint32_t)im.nslots;
    mcus_x = (uint32_t)im.mcus_x;
    one = t.ncomp == 1;
    bw = t.bw ? t.bw : 1;
    by = t.u0 / bw;
    bx = t.u0 - by * bw;
    hs = t.h >= 4 ? 2 : t.h >> 1;
    vs = t.v >= 4 ? 2 : t.v >> 1;
    slot0 = (uint32_t)im.comp_first_slot[t.comps[0]];
    u = t.u0;
    r = 0;
    bpu = t.bpu;
    slots = t.slots;
    cis = t.cis;
  }
  CJD_HD inline int64_t block(int& ci) const {
    if (one) {
      ci = 0;
      const uint32_t mcu = (by >> vs) * mcus_x + (bx >> hs);
      return base + (int64_t)mcu * nslots + slot0 + ((by & ((1u << vs) - 1)) << hs) + (bx & ((1u << hs) - 1));
    }
    ci = (int)((cis >> (2 * r)) & 3u);
    return base + (int64_t)u * nslots + (int64_t)((slots >> (4 * r)) & 15u);
  }
  CJD_HD inline void next() {
    if (one) {
      if (++bx == bw) {
        bx = 0;
        ++by;
      }
    } else if (++r == bpu) {
      r = 0;
      ++u;
    }
  }
};


// Coefficient-buffer index of block a (counted from the scan's first block)
// of a one-component scan, computed on its own (one division).
CJD_HD inline int64_t prog_block_at(const ProgTask& t, const ImageDesc& im, uint32_t a) {
  const uint32_t bw = t.bw ? t.bw : 1, by = a / bw, bx = a - by * bw;
  const uint32_t hs = t.h >= 4 ? 2 : t.h >> 1, vs = t.v >= 4 ? 2 : t.v >> 1;
  const uint32_t mcu = (by >> vs) * (uint32_t)im.mcus_x + (bx >> hs);
  return im.coef_base + (int64_t)mcu * im.nslots + im.comp_first_slot[t.comps[0]] +
         ((by & ((1u << vs) - 1)) << hs) + (bx & ((1u << hs) - 1));
}

// q[0] |= bit by lane 0 (an atomic OR on the word that holds it: a task of the
// same level may store the word's other coefficient at the same time).
CJD_HD inline void prog_dc_or(int16_t* q, int bit) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (__lane_id() == 0)
    __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(q), (uint32_t)bit, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
#else
  q[0] = (int16_t)(q[0] | bit);
#endif
}

// Zigzag positions [a, b] as a mask (0 <= a, b <= 63; empty when a > b).
CJD_HD inline uint64_t prog_span(int a, int b) {
  if (a > b) return 0;
  const uint64_t hi = b >= 63 ? ~0ull : (1ull << (b + 1)) - 1;
  return hi & (~0ull << a);
}

// Refinement state of one block: bit z of nz / neg / has is set where the
// coefficient at zigzag position z is nonzero (within the band) / negative /
// already has bit Al, as the block was when the task reached it (a scan
// changes each coefficient at most once, after passing it).
struct ProgMasks {
  uint64_t nz, neg, has;
};

// The set bits of `mask` whose rank (the number of set bits below) passes
// keep(rank): on the device each lane l ranks its own bit with one popcount
// and a ballot collects the answers; on the host a loop does the same.
template <class F>
CJD_HD inline uint64_t prog_rank_select(uint64_t mask, F&& keep) {
  return prog_ballot([&](int l) {
    return ((mask >> l) & 1) && keep(__builtin_popcountll(mask & ((1ull << l) - 1)));
  });
}

// Correction bits of the nonzero coefficients in `span`, read in position
// order (up to 32 at a time, the i-th bit read going to the i-th position);
// returns the positions whose bit is 1 and that lack bit Al (those move 2^Al
// away from zero).
CJD_HD inline uint64_t prog_corrections(ProgBits& br, const ProgMasks& m, uint64_t span) {
  uint64_t todo = m.nz & span, bump = 0;
  while (todo) {
    const int c = __builtin_popcountll(todo), n = c < 32 ? c : 32;
    const uint64_t part = n == c ? todo : prog_rank_select(todo, [](int r) { return r < 32; });
    const uint32_t bits = br.get(n);  // bit n-1-i belongs to the i-th position of part
    bump |= prog_rank_select(part, [&](int r) { return ((bits >> (n - 1 - r)) & 1u) != 0; });
    todo &= ~part;
  }
  return bump & ~m.has;
}

// One 16-byte-aligned block of 128 bytes held by a lane (the batch fetch).
struct ProgRow {
  uint32_t w[32];
};

// Decode one task (a restart interval of a scan) into coefs; tab holds the
// scan's four Huffman tables (the table set sg.tset; on the device their
// lookups, at wave-uniform addresses, are scalar loads).  DC first: the DC difference as in baseline,
// the predictor stored << Al; DC refine: one raw bit per block ORed in at bit
// Al; AC first: run/size symbols over Ss..Se with values << Al, EOBn (size
// 0, run r < 15) ending this block and the next 2^r + (r extra bits) - 1
// blocks; AC refine: a size-1 symbol's sign bit makes a new coefficient
// +-2^Al at the (run + 1)-th coefficient of the band still zero, every
// nonzero coefficient passed on the way takes a correction bit, ZRL passes 16
// zero coefficients, and EOBn ends the symbols of this block and the next
// blocks of the run, whose nonzero coefficients still take correction bits.
// An invalid code, a refinement size other than 0 or 1, a run past Se, or
// (for an interval followed by an RSTn) data that does not end in the
// interval's last byte sets kStatusSyntax.
CJD_HD inline void prog_task(ProgShared& sh, const uint32_t* words, const ImageDesc* images, const SegmentDesc& sg,
                             const ProgTask& t, const HuffTable* tab, int16_t* coefs, uint32_t* status,
                             uint32_t epoch) {
  const ImageDesc& im = images[sg.image];
  ProgBits br;
  br.init(words + sg.word_base, sg.nwords, sh.ring);
  const uint32_t nb = t.nunits * t.bpu;
  const int p1 = 1 << t.al;
  Lanes<int> nat;  // natural index of the lane's zigzag position
  PROG_LANES(l) nat.at(l) = kDeZigZag[l];
  ProgIter it;
  it.init(t, im);
  bool err = false;
  if (t.ss == 0 && t.ah == 0) {  // DC first
    uint32_t pred[3] = {0, 0, 0};
    for (uint32_t j = 0; j < nb; ++j, it.next()) {
      br.refill();
      int ci;
      const int64_t g = it.block(ci);
      const uint32_t bits = br.peek();
      int len;
      const int sym = huff_sym(tab[ci == 0 ? t.dcs[0] : ci == 1 ? t.dcs[1] : t.dcs[2]], bits, len);
      if (!len) {
        err = true;
        break;
      }
      const int size = sym & 15;
      const uint32_t p = (ci == 0 ? pred[0] : ci == 1 ? pred[1] : pred[2]) + (uint32_t)extend_value(bits, len, size);
      if (ci == 0) pred[0] = p;
      else if (ci == 1) pred[1] = p;
      else pred[2] = p;
      br.skip(len + size);
      PROG_LANES(l) if (l == 0) coefs[g * 64] = (int16_t)(uint16_t)(p << t.al);
    }
  } else if (t.ss == 0) {  // DC refine
    for (uint32_t j = 0; j < nb; ++j, it.next()) {
      br.refill();
      int ci;
      const int64_t g = it.block(ci);
      if (br.get(1)) prog_dc_or(coefs + g * 64, p1);
    }
  } else if (t.ah == 0) {  // AC first
    const HuffTable& act = tab[t.ac];
    uint32_t eobrun = 0;
    Lanes<int16_t> val;
    for (uint32_t j = 0; j < nb && !err; ++j, it.next()) {
      if (eobrun) {
        --eobrun;
        continue;
      }
      br.refill();
      int ci;
      const int64_t g = it.block(ci);
      uint64_t wr = 0;  // positions written in this block
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
          const int16_t v = (int16_t)(uint16_t)((uint32_t)extend_value(bits, len, size) << t.al);
          PROG_LANES(l) if (l == k) val.at(l) = v;
          wr |= 1ull << k;
          ++k;
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
      PROG_LANES(l) if ((wr >> l) & 1) coefs[g * 64 + nat.at(l)] = val.at(l);
    }
  } else {  // AC refine (one component: the scan's own block grid)
    const HuffTable& act = tab[t.ac];
    const uint64_t band = prog_span(t.ss, t.se);
    uint32_t eobrun = 0;
    Lanes<ProgRow> pre;  // the next batch: lane l holds its block
    Lanes<int16_t> val;
    auto fetch = [&](uint32_t b0) {
      PROG_LANES(l) {
        if (b0 + (uint32_t)l < nb) {
          const int16_t* q = coefs + prog_block_at(t, im, t.u0 + b0 + (uint32_t)l) * 64;
#if defined(__HIP_DEVICE_COMPILE__)
          const uint4* s4 = reinterpret_cast<const uint4*>(q);
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint4 v = s4[i];
            pre.at(l).w[4 * i] = v.x;
            pre.at(l).w[4 * i + 1] = v.y;
            pre.at(l).w[4 * i + 2] = v.z;
            pre.at(l).w[4 * i + 3] = v.w;
          }
#else
          memcpy(pre.at(l).w, q, 128);
#endif
        }
      }
    };
    fetch(0);
    for (uint32_t b0 = 0; b0 < nb && !err; b0 += kProgBatch) {
      PROG_SYNC();  // the last batch's LDS reads are done
      PROG_LANES(l) {
        if (b0 + (uint32_t)l < nb) {
#if defined(__HIP_DEVICE_COMPILE__)
          uint4* d4 = reinterpret_cast<uint4*>(sh.blk[l]);
#pragma unroll
          for (int i = 0; i < 8; ++i)
            d4[i] = make_uint4(pre.at(l).w[4 * i], pre.at(l).w[4 * i + 1], pre.at(l).w[4 * i + 2], pre.at(l).w[4 * i + 3]);
#else
          memcpy(sh.blk[l], pre.at(l).w, 128);
#endif
        }
      }
      PROG_SYNC();
      if (b0 + kProgBatch < nb) fetch(b0 + kProgBatch);  // in flight during this batch
      const uint32_t nbat = nb - b0 < (uint32_t)kProgBatch ? nb - b0 : (uint32_t)kProgBatch;
      for (uint32_t i = 0; i < nbat && !err; ++i, it.next()) {
        br.refill();
        int ci;
        const int64_t g = it.block(ci);
        Lanes<int16_t> x;
        PROG_LANES(l) x.at(l) = sh.blk[i][nat.at(l)];
        ProgMasks m;
        m.nz = prog_ballot([&](int l) { return x.at(l) != 0; }) & band;
        m.neg = prog_ballot([&](int l) { return x.at(l) < 0; });
        m.has = prog_ballot([&](int l) { return (x.at(l) & p1) != 0; });
        uint64_t bump = 0, fresh = 0;  // corrected / new coefficients
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
            // the target: the (run + 1)-th coefficient from k on that is
            // zero; the nonzero ones before it take correction bits
            const uint64_t zeros = prog_rank_select(~m.nz & prog_span(k, t.se), [&](int r) { return r == run; });
            if (!zeros) {  // the band ends first
              bump |= prog_corrections(br, m, prog_span(k, t.se));
              k = t.se + 1;
              err = size != 0;
              break;
            }
            const int tz = __builtin_ctzll(zeros);
            bump |= prog_corrections(br, m, prog_span(k, tz - 1));
            if (size) {
              const int16_t nv = (int16_t)(v * p1);
              PROG_LANES(l) if (l == tz) val.at(l) = nv;
              fresh |= 1ull << tz;
            }
            k = tz + 1;
          }
        }
        if (!err && eobrun) {
          bump |= prog_corrections(br, m, prog_span(k, t.se));
          --eobrun;
        }
        PROG_LANES(l) {
          if ((fresh >> l) & 1) {
            coefs[g * 64 + nat.at(l)] = val.at(l);
          } else if ((bump >> l) & 1) {
            const int16_t xv = x.at(l);
            coefs[g * 64 + nat.at(l)] = (int16_t)(xv + (xv >= 0 ? p1 : -p1));
          }
        }
      }
    }
  }
  // an interval followed by an RSTn must end in its last data byte (the
  // serial decoders byte-align there and read the marker)
  if (!err && !t.last && (br.pos + 7) / 8 != sg.nbits / 8) err = true;
  if (err) PROG_LANES(l) if (l == 0) status_or(status + im.error_index, kStatusSyntax, epoch);
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
