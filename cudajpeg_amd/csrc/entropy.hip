// Entropy-decode kernels for the MI355X baseline-JPEG decoder (gfx950).
//
// Replaces the reference's speculative per-bit-position kernels and its host
// serial decode (src/entropyRLEdecodeGPU.cu:7-238, src/decodeScanGPU.cu:123-491,
// src/decodeScanCPU.cu:10-152) with a self-synchronising tile decoder:
//
//  1. The unstuffed segment is cut into tiles of S bits.  For every tile, one
//     lane per MCU slot h decodes speculatively from the tile's first bit as if
//     a block of slot h started there ("hypothesis" h), remembering its first
//     block starts (bit, slot, DC prefix) and its exit: the first block start at
//     or after the tile end, with block count and DC sums.
//  2. Link: from every hypothesis exit of tile t, a lane decodes into tile t+1
//     until it lands on a remembered block start of some hypothesis h' of
//     t+1 (the two trajectories have synchronised), giving a map
//     sigma_t : h -> h' plus the blocks / DC sums gained on the way.
//  3. The maps compose associatively.  Within a workgroup the prefix
//     compositions are built by pointer jumping in LDS; across workgroups a
//     decoupled-lookback scan turns them into the true decoder state at every
//     tile: entry bit, MCU slot, block index and DC predictors.
//  4. Each tile is decoded once more from its true entry and its quantised
//     coefficients are written (natural order, scan order, int16).
//
// The symbol semantics follow the reference bit for bit on valid streams
// (src/decodeScanGPU.cu:50-100): 16-bit peek with 0xFF beyond the end, EXTEND,
// EOB / ZRL, de-zigzag, DC prediction per component, SYNTAX_ERROR on an invalid
// code, a non-ZRL zero-size AC symbol or a coefficient index past 63.
#include <hip/hip_runtime.h>

#include "cjd_internal.h"

namespace cjd {

__constant__ uint8_t kDeZigZag[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15,
    23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// 32 bits of the unstuffed stream starting at bit p, MSB first; bytes beyond
// the segment read as 0xFF (src/decodeScanGPU.cu:32).
__device__ __forceinline__ uint32_t load_word(const uint32_t* __restrict__ w, uint32_t nwords, uint32_t i) {
  return i < nwords ? __builtin_bswap32(w[i]) : 0xFFFFFFFFu;
}
__device__ __forceinline__ uint32_t peek32(const uint32_t* __restrict__ w, uint32_t nwords, uint32_t p) {
  const uint32_t i = p >> 5, o = p & 31;
  const uint64_t v = ((uint64_t)load_word(w, nwords, i) << 32) | load_word(w, nwords, i + 1);
  return (uint32_t)((v << o) >> 32);
}

// Huffman symbol of the code at the top of `bits`; len = code length, or 0 for
// an invalid code (T.81 F.2.2.3 canonical decode for codes longer than the LUT).
__device__ __forceinline__ int huff_sym(const HuffTable& T, uint32_t bits, int& len) {
  const uint32_t e = T.lut[bits >> (32 - kLutBits)];
  len = (int)(e & 31);
  if (len) return (int)((e >> 8) & 0xFF);
  for (int l = kLutBits + 1; l <= 16; ++l) {
    const int32_t code = (int32_t)(bits >> (32 - l));
    if (code <= T.maxcode[l]) {
      len = l;
      return T.vals[(T.valoff[l] + code) & 0xFF];
    }
  }
  len = 0;
  return 0;
}

__device__ __forceinline__ int extend_value(uint32_t bits, int len, int size) {
  if (!size) return 0;
  int v = (int)((bits << len) >> (32 - size));
  if (v < (1 << (size - 1))) v += (-1 << size) + 1;
  return v;
}

// Decode one block at bit p.  kStrict: report errors (err = 1) as the
// reference does; otherwise errors end the block deterministically (garbage
// decoding during speculation).  Writes coefficients to blk when kWrite.
template <bool kWrite>
__device__ __forceinline__ uint32_t decode_block(const uint32_t* __restrict__ w, uint32_t nwords, uint32_t p,
                                                 const HuffTable& dct, const HuffTable& act, int& dcdiff,
                                                 int& err, int16_t* blk) {
  uint32_t bits = peek32(w, nwords, p);
  int len;
  int sym = huff_sym(dct, bits, len);
  if (!len) {  // invalid DC code
    err = 1;
    dcdiff = 0;
    return p + 1;
  }
  int size = sym & 15;
  dcdiff = extend_value(bits, len, size);
  p += len + size;
  int coef = 0;
  while (coef < 63) {
    bits = peek32(w, nwords, p);
    sym = huff_sym(act, bits, len);
    if (!len) { err = 1; return p; }
    size = sym & 15;
    const int v = extend_value(bits, len, size);
    p += len + size;
    if (!sym) break;                           // EOB
    if (!size && sym != 0xF0) { err = 1; return p; }
    coef += (sym >> 4) + 1;
    if (coef > 63) { err = 1; return p; }
    if (kWrite) blk[kDeZigZag[coef]] = (int16_t)v;
  }
  return p;
}

// ---------------------------------------------------------------- shared --
struct EntShared {
  HuffTable tab[4];
  // per-lane hypothesis results
  uint32_t exit_p[kWG];
  uint32_t exit_s[kWG];            // slot
  int32_t tot[kWG][4];             // blocks, dc0, dc1, dc2
  uint32_t nlist[kWG];
  // lists of block starts; reused as coefficient staging in the final pass
  union {
    struct {
      uint32_t p[kWG][kListLen];
      uint32_t s_dc0[kWG][kListLen];   // slot | dc0 << 16
      uint32_t dc12[kWG][kListLen];    // dc1 | dc2 << 16
    } l;
    int16_t coef[kWG][64];
  } u;
  // link maps / prefix compositions (double buffered for pointer jumping)
  int32_t mh[2][kWG];
  int32_t mc[2][kWG][4];
  // lookback result broadcast
  int32_t entry_h;
  int32_t entry_c[4];
  uint32_t item;
};

__device__ __forceinline__ void copy_tables(EntShared& sh, const HuffTable* __restrict__ g) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(g);
  uint32_t* dst = reinterpret_cast<uint32_t*>(sh.tab);
  constexpr int n = (int)(sizeof(HuffTable) * 4 / 4);
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ int slot_comp(const ImageDesc& im, int slot) { return im.slot_comp[slot]; }

// Lookback record helpers (agent-scope acquire/release).
__device__ __forceinline__ int ld_status(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_clear_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ------------------------------------------------------------ main kernel --
__global__ void __launch_bounds__(kWG) k_entropy(const uint32_t* __restrict__ words,
                                                 const ImageDesc* __restrict__ images,
                                                 const SegmentDesc* __restrict__ segs,
                                                 const WorkItem* __restrict__ items, uint32_t nitems,
                                                 const HuffTable* __restrict__ tables,
                                                 int16_t* __restrict__ coefs, uint32_t* __restrict__ status,
                                                 LookbackRec* __restrict__ lb, uint32_t* __restrict__ ticket) {
  __shared__ EntShared sh;
  const int tid = threadIdx.x;
  if (tid == 0) sh.item = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t item_id = sh.item;
  if (item_id >= nitems) return;
  const WorkItem it = items[item_id];
  const SegmentDesc sg = segs[it.segment];
  const ImageDesc& im = images[sg.image];
  const int nh = im.nslots;
  const uint32_t S = sg.tile_bits;
  const uint32_t* __restrict__ w = words + sg.word_base;
  const uint32_t nwords = sg.nwords;
  const int lead = it.t0 > 0 ? 1 : 0;
  const int J = (int)it.ntiles + lead;        // local tiles, lead-in first
  const uint32_t tfirst = it.t0 - lead;       // segment tile index of local tile 0

  copy_tables(sh, tables + im.table_base);
  __syncthreads();

  auto tile_end = [&](uint32_t t) -> uint32_t {
    return (t + 1 >= sg.ntiles) ? sg.nbits : (t + 1) * S;
  };

  // ---- phase 1: speculative decode of every (tile, hypothesis) ----------
  const int j = tid / nh, h = tid - (tid / nh) * nh;
  const bool active = j < J;
  if (active) {
    const uint32_t t = tfirst + (uint32_t)j;
    uint32_t p = t * S;
    int slot = (t == 0) ? 0 : h;
    const uint32_t end = tile_end(t);
    int32_t nb = 0, dc0 = 0, dc1 = 0, dc2 = 0;
    int16_t dummy[1];
    while (p < end) {
      if (nb < kListLen) {
        sh.u.l.p[tid][nb] = p;
        sh.u.l.s_dc0[tid][nb] = (uint32_t)slot | ((uint32_t)(uint16_t)(int16_t)dc0 << 16);
        sh.u.l.dc12[tid][nb] = (uint32_t)(uint16_t)(int16_t)dc1 | ((uint32_t)(uint16_t)(int16_t)dc2 << 16);
      }
      int d, e = 0;
      p = decode_block<false>(w, nwords, p, sh.tab[im.slot_dc[slot]], sh.tab[im.slot_ac[slot]], d, e, dummy);
      const int c = slot_comp(im, slot);
      dc0 += c == 0 ? d : 0;
      dc1 += c == 1 ? d : 0;
      dc2 += c == 2 ? d : 0;
      ++nb;
      if (++slot == nh) slot = 0;
    }
    sh.exit_p[tid] = p;
    sh.exit_s[tid] = (uint32_t)slot;
    sh.tot[tid][0] = nb;
    sh.tot[tid][1] = dc0;
    sh.tot[tid][2] = dc1;
    sh.tot[tid][3] = dc2;
    sh.nlist[tid] = nb < kListLen ? nb : kListLen;
  }
  __syncthreads();

  // ---- phase 2: link every hypothesis exit of tile j to tile j+1 --------
  // sigma in mh[0], counts gained in tile j+1 in mc[0].
  if (active && j + 1 < J) {
    const uint32_t t1 = tfirst + (uint32_t)j + 1;
    const uint32_t end1 = tile_end(t1);
    uint32_t p = sh.exit_p[tid];
    int slot = (int)sh.exit_s[tid];
    int32_t rb = 0, r0 = 0, r1 = 0, r2 = 0;
    int cur[kMaxSlots];
    for (int k = 0; k < nh; ++k) cur[k] = 0;
    int hit = -1, hit_i = 0;
    const int base1 = (j + 1) * nh;
    int16_t dummy[1];
    for (;;) {
      // merge check against the remembered block starts of tile j+1
      for (int k = 0; k < nh && hit < 0; ++k) {
        const int L = base1 + k;
        const int n = (int)sh.nlist[L];
        int c = cur[k];
        while (c < n && sh.u.l.p[L][c] < p) ++c;
        cur[k] = c;
        if (c < n && sh.u.l.p[L][c] == p && (int)(sh.u.l.s_dc0[L][c] & 0xFF) == slot) { hit = k; hit_i = c; }
      }
      if (hit >= 0 || p >= end1) break;
      int d, e = 0;
      p = decode_block<false>(w, nwords, p, sh.tab[im.slot_dc[slot]], sh.tab[im.slot_ac[slot]], d, e, dummy);
      const int c = slot_comp(im, slot);
      r0 += c == 0 ? d : 0;
      r1 += c == 1 ? d : 0;
      r2 += c == 2 ? d : 0;
      ++rb;
      if (++slot == nh) slot = 0;
    }
    int sig = -1;
    int32_t c4[4] = {rb, r0, r1, r2};
    if (hit >= 0) {
      const int L = base1 + hit;
      const uint32_t sd = sh.u.l.s_dc0[L][hit_i], d12 = sh.u.l.dc12[L][hit_i];
      sig = hit;
      c4[0] += sh.tot[L][0] - hit_i;
      c4[1] += sh.tot[L][1] - (int32_t)(int16_t)(sd >> 16);
      c4[2] += sh.tot[L][2] - (int32_t)(int16_t)(d12 & 0xFFFF);
      c4[3] += sh.tot[L][3] - (int32_t)(int16_t)(d12 >> 16);
    } else {
      for (int k = 0; k < nh; ++k) {
        const int L = base1 + k;
        if (sh.exit_p[L] == p && (int)sh.exit_s[L] == slot) { sig = k; break; }
      }
    }
    sh.mh[0][tid] = sig;
    for (int q = 0; q < 4; ++q) sh.mc[0][tid][q] = c4[q];
  }
  __syncthreads();

  // ---- phase 3: prefix compositions by pointer jumping ------------------
  // G_j(h) = state in local tile j when the lead-in follows hypothesis h.
  // Step input F_j = tau_{j-1} (identity for j = 0).
  int buf = 1;
  if (active) {
    if (j == 0) {
      sh.mh[1][tid] = h;
      for (int q = 0; q < 4; ++q) sh.mc[1][tid][q] = 0;
    } else {
      const int L = (j - 1) * nh + h;
      sh.mh[1][tid] = sh.mh[0][L];
      for (int q = 0; q < 4; ++q) sh.mc[1][tid][q] = sh.mc[0][L][q];
    }
  }
  __syncthreads();
  for (int d = 1; d < J; d <<= 1) {
    const int nb = buf ^ 1;
    if (active) {
      if (j >= d) {
        // new G_j(h) = G_j(G_{j-d}(h))
        const int L0 = (j - d) * nh + h;
        const int h1 = sh.mh[buf][L0];
        if (h1 < 0) {
          sh.mh[nb][tid] = -1;
          for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = 0;
        } else {
          const int L1 = j * nh + h1;
          sh.mh[nb][tid] = sh.mh[buf][L1];
          for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = sh.mc[buf][L0][q] + sh.mc[buf][L1][q];
        }
      } else {
        sh.mh[nb][tid] = sh.mh[buf][tid];
        for (int q = 0; q < 4; ++q) sh.mc[nb][tid][q] = sh.mc[buf][tid][q];
      }
    }
    __syncthreads();
    buf = nb;
  }

  // ---- phase 4: decoupled lookback across workgroups --------------------
  if (tid == 0) {
    LookbackRec* me = lb + item_id;
    const int Lq = (J - 1) * nh;
    int eh;
    int32_t ec[4];
    if (!lead) {
      // segment start: tile 0 is decoded from the known state by hypothesis 0
      eh = 0;
      for (int q = 0; q < 4; ++q) ec[q] = sh.tot[0][q];
    } else {
      for (int k = 0; k < nh; ++k) {
        me->agg_h[k] = (int8_t)sh.mh[buf][Lq + k];
        for (int q = 0; q < 4; ++q) me->agg_c[k][q] = sh.mc[buf][Lq + k][q];
      }
      st_status(&me->status, 1);
      // acc = composition of the aggregates seen so far (identity first)
      int acc_h[kMaxSlots];
      int32_t acc_c[kMaxSlots][4];
      for (int k = 0; k < nh; ++k) {
        acc_h[k] = k;
        for (int q = 0; q < 4; ++q) acc_c[k][q] = 0;
      }
      uint32_t q_item = it.prev_item;
      for (;;) {
        LookbackRec* r = lb + q_item;
        int st;
        while ((st = ld_status(&r->status)) == 0) __builtin_amdgcn_s_sleep(1);
        if (st == 2) {
          const int hb = r->incl_h;
          if (hb < 0) {
            eh = -1;
            for (int q = 0; q < 4; ++q) ec[q] = 0;
          } else {
            eh = acc_h[hb];
            for (int q = 0; q < 4; ++q) ec[q] = r->incl_c[q] + acc_c[hb][q];
          }
          break;
        }
        // acc <- acc o agg_q
        int nh_[kMaxSlots];
        int32_t nc[kMaxSlots][4];
        for (int k = 0; k < nh; ++k) {
          const int h1 = r->agg_h[k];
          if (h1 < 0 || acc_h[h1] < 0) {
            nh_[k] = -1;
            for (int q = 0; q < 4; ++q) nc[k][q] = 0;
          } else {
            nh_[k] = acc_h[h1];
            for (int q = 0; q < 4; ++q) nc[k][q] = r->agg_c[k][q] + acc_c[h1][q];
          }
        }
        for (int k = 0; k < nh; ++k) {
          acc_h[k] = nh_[k];
          for (int q = 0; q < 4; ++q) acc_c[k][q] = nc[k][q];
        }
        q_item = items[q_item].prev_item;
      }
    }
    // inclusive state at the last local tile
    int ih = -1;
    int32_t ic[4] = {0, 0, 0, 0};
    if (eh >= 0) {
      const int L = Lq + eh;
      ih = sh.mh[buf][L];
      for (int q = 0; q < 4; ++q) ic[q] = ec[q] + sh.mc[buf][L][q];
    }
    me->incl_h = ih;
    for (int q = 0; q < 4; ++q) me->incl_c[q] = ic[q];
    st_status(&me->status, 2);
    sh.entry_h = eh;
    for (int q = 0; q < 4; ++q) sh.entry_c[q] = ec[q];
  }
  __syncthreads();

  // ---- phase 5: final decode of each tile from its true entry -----------
  const int eh = sh.entry_h;
  const int fj = tid + lead;  // local tile decoded by this lane
  if (fj < J) {
    const uint32_t t = tfirst + (uint32_t)fj;
    uint32_t* st = status + im.error_index;
    if (eh < 0) {
      atomicOr(st, kStatusUnresolved);
    } else {
      // state at tile fj-1 (entry) and at tile fj (exit)
      int hp, hx;
      int32_t cp[4], cx[4];
      uint32_t p;
      int slot;
      if (!lead && fj == 0) {
        hp = 0;
        for (int q = 0; q < 4; ++q) cp[q] = 0;
        p = 0;
        slot = 0;
      } else {
        const int Lp = (fj - 1) * nh + eh;
        hp = sh.mh[buf][Lp];
        for (int q = 0; q < 4; ++q) cp[q] = sh.entry_c[q] + sh.mc[buf][Lp][q];
        if (hp >= 0) {
          p = sh.exit_p[(fj - 1) * nh + hp];
          slot = (int)sh.exit_s[(fj - 1) * nh + hp];
        } else {
          p = 0;
          slot = 0;
        }
      }
      if (!lead && fj == 0) {
        hx = 0;
        for (int q = 0; q < 4; ++q) cx[q] = sh.tot[0][q];
      } else {
        const int Lx = fj * nh + eh;
        hx = sh.mh[buf][Lx];
        for (int q = 0; q < 4; ++q) cx[q] = sh.entry_c[q] + sh.mc[buf][Lx][q];
      }
      if (hp < 0 || hx < 0) {
        atomicOr(st, kStatusUnresolved);
      } else {
        const uint32_t end = tile_end(t);
        int64_t bidx = cp[0];
        int dc[3] = {cp[1], cp[2], cp[3]};
        int16_t* blk = sh.u.coef[tid];
        int err = 0;
        int nb = 0;
        bool stopped = false;
        while (p < end) {
          if (bidx >= (int64_t)sg.nblocks) { stopped = true; break; }
          uint4* b4 = reinterpret_cast<uint4*>(blk);
          for (int k = 0; k < 8; ++k) b4[k] = make_uint4(0, 0, 0, 0);
          int d;
          p = decode_block<true>(w, nwords, p, sh.tab[im.slot_dc[slot]], sh.tab[im.slot_ac[slot]], d, err, blk);
          if (err) break;
          const int c = slot_comp(im, slot);
          dc[c] += d;
          blk[0] = (int16_t)dc[c];
          uint4* g = reinterpret_cast<uint4*>(coefs + (sg.first_block + bidx) * 64);
          for (int k = 0; k < 8; ++k) g[k] = b4[k];
          ++bidx;
          ++nb;
          if (++slot == nh) slot = 0;
        }
        if (err) {
          atomicOr(st, kStatusSyntax);
        } else if (!stopped) {
          const int Lx = fj * nh + hx;
          if (p != sh.exit_p[Lx] || (uint32_t)slot != sh.exit_s[Lx] || bidx != (int64_t)cx[0] ||
              dc[0] != cx[1] || dc[1] != cx[2] || dc[2] != cx[3])
            atomicOr(st, kStatusInconsistent);
          if (t + 1 == sg.ntiles && bidx < (int64_t)sg.nblocks) atomicOr(st, kStatusTruncated);
        }
      }
    }
  }
}

// Serial fallback: one lane decodes a whole segment from its known start, in
// the reference's order (src/decodeScanGPU.cu:435-451).  Used when a tile link
// could not be resolved, and as a plain cross-check.
__global__ void k_entropy_serial(const uint32_t* __restrict__ words, const ImageDesc* __restrict__ images,
                                 const SegmentDesc* __restrict__ segs, uint32_t nsegs,
                                 const HuffTable* __restrict__ tables, int16_t* __restrict__ coefs,
                                 uint32_t* __restrict__ status) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsegs) return;
  const SegmentDesc sg = segs[s];
  const ImageDesc& im = images[sg.image];
  const HuffTable* tab = tables + im.table_base;
  const uint32_t* w = words + sg.word_base;
  uint32_t p = 0;
  int slot = 0, dc[3] = {0, 0, 0};
  int16_t blk[64];
  for (uint32_t b = 0; b < sg.nblocks; ++b) {
    for (int k = 0; k < 64; ++k) blk[k] = 0;
    int d, err = 0;
    p = decode_block<true>(w, sg.nwords, p, tab[im.slot_dc[slot]], tab[im.slot_ac[slot]], d, err, blk);
    if (err) {
      atomicOr(status + im.error_index, kStatusSyntax);
      return;
    }
    const int c = im.slot_comp[slot];
    dc[c] += d;
    blk[0] = (int16_t)dc[c];
    int16_t* g = coefs + (sg.first_block + b) * 64;
    for (int k = 0; k < 64; ++k) g[k] = blk[k];
    if (++slot == im.nslots) slot = 0;
  }
}

}  // namespace cjd
