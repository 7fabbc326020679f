// Host-side launchers of the HIP kernels (defined beside the kernels).
#pragma once
#include <hip/hip_runtime.h>

#include "cjd_internal.h"

namespace cjd {

// Buffers and launch constants of the tiled entropy decoder.
struct EntArgs {
  const uint32_t* words;
  const ImageDesc* images;
  const SegmentDesc* segs;
  const WorkItem* items;
  const HuffTable* tables;
  const uint16_t* luts;
  uint32_t* ebuf;
  BlockRec* recs;
  uint32_t* status;
  LookbackRec* lb;
  unsigned long long* done;  // finished work items (read by the host)
  uint32_t epoch;
  uint64_t wait_ticks;
};

// Buffers of the pixel stage.
struct PixArgs {
  const ImageDesc* images;
  const PixItem* items;
  const BlockRec* recs;
  const uint32_t* ebuf;
  uint64_t ewords;
  const int16_t* coefs;
  const uint16_t* qtabs;
  const uint4* btab;
  uint8_t* out;
};

struct UnstuffChunk;
// Fill the batch's scan words with 0xFF, then unstuff every chunk of stuffed
// scan bytes into them (unstuff.hip).
hipError_t launch_unstuff(const UnstuffChunk* chunks, uint32_t nchunks, const uint8_t* raw, const SegmentDesc* segs,
                          uint32_t* words, uint64_t nwords, hipStream_t s);

hipError_t launch_clear(uint32_t* p, uint64_t nwords, hipStream_t s);
// First kernel of a captured decode (entropy.hip: k_decode_begin).
hipError_t launch_decode_begin(uint32_t* scratch, uint32_t nstatus, uint32_t nrec, hipStream_t s);
// Tiled entropy decode of work items [item0, item_end) (whole segments).
hipError_t launch_entropy(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                          const WorkItem* items, uint32_t item0, uint32_t item_end, const HuffTable* tables,
                          const uint16_t* luts, uint32_t* ebuf, BlockRec* recs, uint32_t* status, LookbackRec* lb,
                          unsigned long long* ticket, unsigned long long* done, uint64_t ticket_base, uint32_t epoch,
                          uint64_t wait_ticks, hipStream_t s);
hipError_t launch_expand(const BlockRec* recs, const uint32_t* ebuf, uint64_t ewords, int16_t* coefs, uint64_t nblocks,
                         hipStream_t s);
hipError_t launch_entropy_serial(const uint32_t* words, const ImageDesc* images, const SegmentDesc* segs,
                                 uint32_t nsegs, const HuffTable* tables, int16_t* coefs, BlockRec* recs,
                                 uint32_t* status, uint32_t epoch, hipStream_t s);
// MFMA table operands of every image (kMfmaTabVecs uint4 per image).
hipError_t launch_mfma_tables(const ImageDesc* images, uint32_t nimages, const uint16_t* qtabs, uint4* btab,
                              hipStream_t s);
// Progressive images (CJD_EXT_PROGRESSIVE): k_prog_prepare over their
// prog_blocks blocks, then one k_prog launch per level (tasks
// [plevel[l], plevel[l + 1]) of the host array plevel, nlevels + 1 entries).
hipError_t launch_prog(const ProgImage* pimgs, uint32_t npimgs, int64_t prog_blocks, const ProgTask* tasks,
                       const uint32_t* plevel, uint32_t nlevels, const uint32_t* words, const ImageDesc* images,
                       const SegmentDesc* segs, const HuffTable* tables, int16_t* coefs, BlockRec* recs,
                       uint32_t* status, uint32_t epoch, hipStream_t s);
hipError_t launch_pixel(int mode, const ImageDesc* images, const PixItem* items, uint32_t nitems,
                        const BlockRec* recs, const uint32_t* ebuf, uint64_t ewords, const int16_t* coefs,
                        const uint16_t* qtabs, const uint4* btab, uint8_t* out, hipStream_t s);

}  // namespace cjd
