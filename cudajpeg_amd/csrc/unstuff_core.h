// Device-side unstuffing of entropy-coded scans (the work of
// src/decodeScanGPU.cu:11-24, done by the host planner otherwise): remove the
// 0x00 after every 0xFF, keep RSTn (or split the scan into restart segments
// at it), keep the 0xFF and drop the second byte of any other pair.
//
// The scan's stuffed bytes are cut into chunks of kUnstuffChunk bytes, one
// 256-lane workgroup each, 16 bytes per lane.  Whether a byte is the second
// byte of a pair depends on every 0xFF before it, so the host planner, which
// walks the scan's 0xFF bytes anyway to size the segments, gives each chunk
// its starting state (UnstuffChunk); within a chunk each lane summarises its
// 16 bytes as a function "is my first byte consumed -> is the next lane's",
// the functions are composed over the lanes by a scan, and a second scan
// over the lanes' output counts (segmented at RSTn splits) gives every
// output byte its place.  The phases are __host__ __device__ functions over
// the workgroup's shared state, so the CPU emulator runs the same code lane
// by lane with barriers between phases.
#pragma once
#include <stdint.h>

#include "cjd_internal.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CJD_UHD __host__ __device__
#else
#define CJD_UHD
#endif

namespace cjd {

constexpr int kUnstuffLanes = 256;
constexpr int kUnstuffLaneBytes = 16;
constexpr int kUnstuffChunk = kUnstuffLanes * kUnstuffLaneBytes;

enum : uint32_t {
  kChunkSkipFirst = 1u,  // the first byte is the second byte of a pair begun in the previous chunk
  kChunkSplit = 2u,      // RSTn ends a segment (the image has a restart interval)
  kChunkScanEnd = 4u,    // no byte follows the chunk in its scan
};

// Starting state of one chunk.
struct UnstuffChunk {
  uint64_t src;    // byte offset of the chunk's first byte in the raw scan buffer
  uint32_t len;    // stuffed bytes in the chunk (1 .. kUnstuffChunk)
  uint32_t seg;    // global segment the chunk's first output byte belongs to
  uint32_t dst;    // byte offset of that output byte within its segment
  uint32_t flags;  // kChunk*
};

struct UnstuffShared {
  uint8_t fn[kUnstuffLanes];         // carry function of each lane, then the composed prefix
  uint32_t brk[kUnstuffLanes];       // RSTn splits in the lane (then inclusive prefix)
  uint32_t tail[kUnstuffLanes];      // output bytes after the lane's last split, or all when none (then prefix)
  alignas(16) uint8_t bytes[kUnstuffChunk + 16];  // the chunk, then the byte after it (0xFF at the scan's end)
  // a chunk with no RSTn split: its output bytes gathered here (the chunk's
  // output starts at byte `lead` of the buffer, the alignment of its first
  // destination byte), then written out as aligned words
  alignas(16) uint8_t outb[kUnstuffChunk + 16];
};

// Phase 0 (lane): the chunk's bytes into LDS, 16 per lane; lane 0 adds the
// byte that follows the chunk.  The raw buffer holds each scan at a 16-byte
// aligned offset and chunks start 4096 bytes apart within it, so on the
// device a lane's 16 bytes are one aligned 16-byte load (the last lane's may
// run past the chunk into the scan's padding; those bytes are zeroed).
CJD_UHD inline void unstuff_load(UnstuffShared& sh, int lane, const UnstuffChunk& c, const uint8_t* raw) {
  const uint32_t k0 = (uint32_t)(lane * kUnstuffLaneBytes);
#if defined(__HIP_DEVICE_COMPILE__)
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  if (k0 < c.len) v = *reinterpret_cast<const uint4*>(raw + c.src + k0);
  if (k0 + kUnstuffLaneBytes > c.len) {  // zero the bytes past the chunk
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (int i = 0; i < 16; ++i)
      if (k0 + (uint32_t)i >= c.len) w[i >> 2] &= ~(0xFFu << (8 * (i & 3)));
    v = make_uint4(w[0], w[1], w[2], w[3]);
  }
  *reinterpret_cast<uint4*>(&sh.bytes[k0]) = v;
#else
  for (int i = 0; i < kUnstuffLaneBytes; ++i) {
    const uint32_t k = k0 + (uint32_t)i;
    sh.bytes[k] = k < c.len ? raw[c.src + k] : 0u;
  }
#endif
}

// After phase 0's barrier (lane 0): the byte that follows the chunk.
CJD_UHD inline void unstuff_load_next(UnstuffShared& sh, const UnstuffChunk& c, const uint8_t* raw) {
  sh.bytes[c.len] = (c.flags & kChunkScanEnd) ? 0xFFu : raw[c.src + c.len];
}

// Whether byte k of the chunk is a 0xFF that begins a pair, given whether it
// is consumed as the second byte of the pair before it.
CJD_UHD inline bool unstuff_starts_pair(const UnstuffShared& sh, uint32_t k, bool consumed) {
  return !consumed && sh.bytes[k] == 0xFFu;
}

// Phase 1 (lane): the lane's carry function, bit x = whether the next lane's
// first byte is consumed when this lane's first byte is (x = 1) or is not.
CJD_UHD inline void unstuff_carry(UnstuffShared& sh, int lane, const UnstuffChunk& c) {
  uint8_t f = 0;
  const uint32_t k0 = (uint32_t)(lane * kUnstuffLaneBytes);
  for (int x = 0; x < 2; ++x) {
    bool consumed = x != 0;
    for (uint32_t k = k0; k < k0 + kUnstuffLaneBytes; ++k) {
      if (k >= c.len) {  // past the chunk: nothing is consumed after it
        consumed = false;
        break;
      }
      consumed = unstuff_starts_pair(sh, k, consumed);
    }
    f |= (uint8_t)((consumed ? 1u : 0u) << x);
  }
  sh.fn[lane] = f;
}

// h = g after f, on two-entry tables of bits.
CJD_UHD inline uint8_t compose_carry(uint8_t g, uint8_t f) {
  return (uint8_t)(((g >> (f & 1u)) & 1u) | (((g >> ((f >> 1) & 1u)) & 1u) << 1));
}

// Inclusive Hillis-Steele step d of the function scan (all lanes read, then
// all write: two barriers per step on the device).
CJD_UHD inline uint8_t unstuff_fn_read(const UnstuffShared& sh, int lane, int d) {
  return lane >= d ? compose_carry(sh.fn[lane], sh.fn[lane - d]) : sh.fn[lane];
}

// Whether lane `lane`'s first byte is consumed (after the scan of fn).
CJD_UHD inline bool unstuff_carry_in(const UnstuffShared& sh, int lane, const UnstuffChunk& c) {
  const bool skip = (c.flags & kChunkSkipFirst) != 0;
  if (lane == 0) return skip;
  return ((sh.fn[lane - 1] >> (skip ? 1 : 0)) & 1u) != 0;
}

// Walk the lane's bytes: emit(byte) for every output byte and split() at
// every RSTn that ends a segment.
template <class Emit, class Split>
CJD_UHD inline void unstuff_walk_lane(const UnstuffShared& sh, int lane, const UnstuffChunk& c, bool consumed,
                                      Emit&& emit, Split&& split) {
  const uint32_t k0 = (uint32_t)(lane * kUnstuffLaneBytes);
  for (uint32_t k = k0; k < k0 + kUnstuffLaneBytes && k < c.len; ++k) {
    if (consumed) {
      consumed = false;
      continue;
    }
    const uint8_t v = sh.bytes[k];
    if (v != 0xFFu) {
      emit(v);
      continue;
    }
    consumed = true;
    const uint8_t nx = sh.bytes[k + 1];  // the byte after the chunk sits at bytes[len]
    if (nx == 0x00u) {
      emit(0xFFu);
    } else if ((nx & 0xF8u) == 0xD0u) {
      if (c.flags & kChunkSplit) {
        split();
      } else {
        emit(0xFFu);
        emit(nx);
      }
    } else {
      emit(0xFFu);
    }
  }
}

// Phase 2 (lane): the lane's splits and its output bytes after the last one.
CJD_UHD inline void unstuff_count(UnstuffShared& sh, int lane, const UnstuffChunk& c, bool consumed) {
  uint32_t brk = 0, tail = 0;
  unstuff_walk_lane(
      sh, lane, c, consumed, [&](uint8_t) { ++tail; },
      [&]() {
        ++brk;
        tail = 0;
      });
  sh.brk[lane] = brk;
  sh.tail[lane] = tail;
}

// Step d of the segmented scan of (splits, bytes since the last split).
CJD_UHD inline void unstuff_count_read(const UnstuffShared& sh, int lane, int d, uint32_t& brk, uint32_t& tail) {
  brk = sh.brk[lane];
  tail = sh.tail[lane];
  if (lane >= d) {
    const uint32_t pb = sh.brk[lane - d], pt = sh.tail[lane - d];
    tail = brk ? tail : pt + tail;
    brk += pb;
  }
}

// Phase 3 (lane): write the lane's output bytes, out[i] = byte i of the word
// buffer.  seg_base(s) = byte offset of segment s's data in the word buffer;
// it is asked for only when a byte goes to segment s (an RSTn that ends the
// scan opens no segment).
template <class Out, class SegBase>
CJD_UHD inline void unstuff_write(const UnstuffShared& sh, int lane, const UnstuffChunk& c, bool consumed,
                                  Out&& out, SegBase&& seg_base) {
  uint32_t seg = c.seg, pos = c.dst;
  if (lane > 0) {
    const uint32_t pb = sh.brk[lane - 1], pt = sh.tail[lane - 1];
    seg += pb;
    pos = pb ? pt : c.dst + pt;
  }
  uint64_t base = 0;
  bool have_base = false;
  unstuff_walk_lane(
      sh, lane, c, consumed,
      [&](uint8_t v) {
        if (!have_base) {
          base = seg_base(seg);
          have_base = true;
        }
        out[base + pos++] = v;
      },
      [&]() {
        ++seg;
        pos = 0;
        have_base = false;
      });
}

// Whether the chunk splits at no RSTn (after the count scan): its output is
// then one run of sh.tail[last] bytes from (c.seg, c.dst).
CJD_UHD inline bool unstuff_single_run(const UnstuffShared& sh) { return sh.brk[kUnstuffLanes - 1] == 0; }

// Phase 3 of a chunk without splits (lane): the lane's output bytes into
// sh.outb at lead + their place in the run (lead = the run's destination
// byte address mod 16).
CJD_UHD inline void unstuff_gather(UnstuffShared& sh, int lane, const UnstuffChunk& c, bool consumed, uint32_t lead) {
  uint32_t pos = lead + (lane > 0 ? sh.tail[lane - 1] : 0u);
  unstuff_walk_lane(
      sh, lane, c, consumed, [&](uint8_t v) { sh.outb[pos++] = v; }, [&]() {});
}

// Phase 4 (lane, after a barrier): the run written from sh.outb, aligned
// 16-byte words where the whole word is the run's, bytes at the two ends.
// dst = the run's first destination byte (16-byte aligned dst - lead).
template <class Out>
CJD_UHD inline void unstuff_flush(const UnstuffShared& sh, int lane, uint32_t n, uint32_t lead, uint64_t dst0,
                                  Out&& out) {
  // bytes [lead, lead + n) of outb go to out[dst0 - lead + j] for j in that range
  const uint64_t base = dst0 - lead;  // 16-byte aligned
  const uint32_t end = lead + n;
  for (uint32_t q = (uint32_t)lane; q * 16u < end; q += kUnstuffLanes) {
    const uint32_t j0 = q * 16u;
    if (j0 >= lead && j0 + 16u <= end) {
#if defined(__HIP_DEVICE_COMPILE__)
      *reinterpret_cast<uint4*>(&out[base + j0]) = *reinterpret_cast<const uint4*>(&sh.outb[j0]);
#else
      for (uint32_t j = j0; j < j0 + 16u; ++j) out[base + j] = sh.outb[j];
#endif
    } else {
      for (uint32_t j = j0; j < j0 + 16u; ++j)
        if (j >= lead && j < end) out[base + j] = sh.outb[j];
    }
  }
}

}  // namespace cjd
