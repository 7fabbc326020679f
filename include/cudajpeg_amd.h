/*
 * cudajpeg_amd -- C ABI of the MI355X-native baseline-JPEG decoder.
 *
 * Plain pointers and sizes only; no C++ types, no exceptions, no torch types.
 * Return values use the error numbering of the reference (include/format.h:7-14).
 * Device memory is owned by the context/batch; host buffers by the caller.
 * Streams cross as opaque `void*` (a hipStream_t; NULL = the null stream).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef CUDAJPEG_AMD_H
#define CUDAJPEG_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/format.h:7-14 */
#define CJD_NO_ERROR 0
#define CJD_SYNTAX_ERROR 1
#define CJD_UNSUPPORTED_ERROR 2
#define CJD_OOM_ERROR 3
#define CJD_DEVICE_MEM_ERROR 4      /* reference CUDA_MEM_ERROR */
#define CJD_FILE_ERROR 5
#define CJD_KERNEL_LAUNCH_ERROR 6   /* reference CUDA_KERNEL_LAUNCH_ERROR */
#define CJD_PROGRAMMER_ERROR (-1)

/* iDCT variants of the fused pixel stage */
#define CJD_IDCT_INT 0    /* reference integer Chen-Wang iDCT, bit-exact pixels */
#define CJD_IDCT_FLOAT 1  /* FP32 VALU separable iDCT */
#define CJD_IDCT_MFMA 2   /* FP16 hi/lo 64x64 contraction on the matrix cores, FP32 accumulate */

/* entropy decoders */
#define CJD_ENTROPY_TILED 0   /* self-synchronising tiles + lookback scan (default) */
#define CJD_ENTROPY_SERIAL 1  /* one lane per restart segment (reference order) */

/* what cjd_batch_download copies */
#define CJD_OUT_PIXELS 0        /* interleaved RGB (W*H*3) or gray (W*H) */
#define CJD_OUT_COEFFICIENTS 1  /* int16 [num_blocks][64], quantised, natural order, scan order
                                   (non-interleaved files: the interleaved scan's order) */
#define CJD_OUT_PLANE0 2        /* planar component c = CJD_OUT_PLANE0 + c, stride x plane height */

/* Frame description produced by the host parser: what decodeSOF / decodeDHT /
 * decodeDQT / decodeDRI and the SOS header leave in the reference JPGReader
 * (src/format.cu:242-442, src/decodeScanGPU.cu:124-141). */
typedef struct {
  int32_t width, height, num_channels;
  int32_t mcus_x, mcus_y;               /* reference num_blocks_x / num_blocks_y */
  int32_t mcu_width, mcu_height;        /* reference block_size_x / block_size_y */
  int32_t restart_interval;
  int32_t blocks_per_mcu;
  int64_t num_blocks;
  int32_t comp_id[3];
  int32_t samples_x[3], samples_y[3];
  int32_t comp_width[3], comp_height[3];
  int32_t comp_stride[3], comp_plane_height[3];
  int32_t dq_id[3], dc_id[3], ac_id[3]; /* ac_id is 2|Ta as in the reference */
  uint8_t dq_tables[4][64];             /* as stored in the file: zigzag order */
  uint8_t huff_bits[4][16];             /* table id = Tc*2 + Th */
  uint8_t huff_vals[4][256];
  uint64_t scan_offset;                 /* entropy-coded segment: file[scan_offset, +scan_length) */
  uint64_t scan_length;
  int32_t error_line;                   /* source line of the failed check, 0 if none */
  int32_t num_scans;                    /* 1, or with CJD_EXT_NONINTERLEAVED the one-component scans,
                                           with CJD_EXT_PROGRESSIVE the progressive scans
                                           (scan_offset / scan_length: the first) */
  uint16_t dq_tables16[4][64];          /* the same tables with 16-bit entries (CJD_EXT_DQT16), zigzag */
  int32_t progressive;                  /* 1 for a progressive frame (CJD_EXT_PROGRESSIVE) */
  int32_t reserved_;
} cjd_frame_info;

/* Extensions beyond the reference decoder (off by default: the reference's
 * behaviour and error codes):
 *   CJD_EXT_DQT16  16-bit quantisation tables (DQT Pq = 1, which
 *                  src/format.cu:432 rejects with SYNTAX_ERROR); dq_tables
 *                  then holds the entries clamped to 255, dq_tables16 the
 *                  values. */
#define CJD_EXT_DQT16 1
/*   CJD_EXT_NONINTERLEAVED  non-interleaved baseline scans: one scan per
 *                  component (T.81 A.2.2-A.2.3), which src/decodeScanGPU.cu
 *                  rejects (:124-131).  Scans may be separated by DHT, DQT
 *                  and DRI; each uses the Huffman tables and restart
 *                  interval in effect at its SOS, and each component the
 *                  quant table in effect when its scan begins.  Blocks and
 *                  coefficients are laid out as the interleaved scan would
 *                  order them, with zero blocks where MCUs pad a component's
 *                  own block grid (outside the image). */
#define CJD_EXT_NONINTERLEAVED 2
/*   CJD_EXT_PROGRESSIVE  progressive frames (SOF2, T.81 G.1), which
 *                  src/format.cu:198-200 rejects as an unknown marker with
 *                  SYNTAX_ERROR: DC first and refinement scans (interleaved
 *                  or not), AC first scans with end-of-band runs and AC
 *                  refinement scans, each with the Huffman tables and restart
 *                  interval in effect at its SOS; each component is
 *                  dequantised with the quant table in effect at its first
 *                  scan.  The progression must follow T.81 G.1.1.1 (a band's
 *                  first scan has Ah = 0, each later one Ah = the previous
 *                  Al and Al = Ah - 1, Al <= 13, a component's AC scans after
 *                  its DC first scan); anything else is SYNTAX_ERROR.  The
 *                  scans are decoded on the GPU one lane per restart interval
 *                  of a scan, scans that code disjoint coefficients side by
 *                  side; coefficients come out in the interleaved layout as
 *                  for CJD_EXT_NONINTERLEAVED. */
#define CJD_EXT_PROGRESSIVE 4

typedef struct cjd_context cjd_context;
typedef struct cjd_batch cjd_batch;

/* Marker, quant-table and Huffman-table parsing of an in-memory file
 * (replaces openJPG's marker loop, src/format.cu:113-215, with decodeSOF
 * :242-354, decodeDHT :357-400, decodeDRI :404-421, decodeDQT :424-442 and the
 * SOS header check of src/decodeScanGPU.cu:124-141).  Stops at the scan. */
int cjd_parse(const uint8_t *data, size_t size, cjd_frame_info *out);
/* cjd_parse with the extensions `ext` (CJD_EXT_*) accepted. */
int cjd_parse_ext(const uint8_t *data, size_t size, int ext, cjd_frame_info *out);

/* Device context (replaces the device half of newJPGReader / delJPGReader,
 * src/format.cu:12-59). */
int cjd_context_create(int device, cjd_context **out);
void cjd_context_destroy(cjd_context *ctx);

/* A batch of images decoded together: device-resident scans, tables, work
 * lists, coefficients and frames (replaces the per-reader device buffers
 * managed by ensureMemSize, src/format.cu:81-110). */
int cjd_batch_create(cjd_context *ctx, cjd_batch **out);
void cjd_batch_destroy(cjd_batch *b);

/* Options: iDCT variant, entropy decoder, tile length in bits (512 .. 8192,
 * a multiple of 32, 0 = default 4096), whether planar component planes are
 * also written. */
int cjd_batch_set_options(cjd_batch *b, int idct_mode, int entropy_mode, int tile_bits, int want_planes);

/* Tiled entropy decoder tuning (no counterpart in the reference, whose scan
 * decoder has no tiles): the spacing in bits of the checkpoints where a
 * tile's link can join its speculative pass (>= 64, 0 = tile length / 16) and
 * tiles per work item (2 .. 256, 0 = 256).  Takes effect at the next
 * cjd_batch_stage. */
int cjd_batch_set_tuning(cjd_batch *b, int ck_bits, int item_tiles);

/* Extensions (CJD_EXT_*) the batch accepts in the files it stages; 0 (the
 * default) parses as the reference does.  From the next stage. */
int cjd_batch_set_extensions(cjd_batch *b, int ext);

/* Where the scans are unstuffed (the 0x00 after each 0xFF removed and the
 * scan split at RSTn: src/decodeScanGPU.cu:11-24, which the reference does on
 * the host): CJD_UNSTUFF_DEVICE (default) uploads each file's stuffed scan
 * and unstuffs it with a kernel on the staging stream; CJD_UNSTUFF_HOST
 * unstuffs in the host planner and uploads the words.  Results are identical.
 * Applies from the next stage.  No counterpart in the reference. */
#define CJD_UNSTUFF_HOST 0
#define CJD_UNSTUFF_DEVICE 1
int cjd_batch_set_unstuff(cjd_batch *b, int mode);

/* Decode in parts (no counterpart in the reference; 2 parts by default): for
 * a batch of at least 64 entropy work items, cjd_batch_decode cuts the images
 * into `parts` runs of about equal work (whole images) and launches the runs'
 * entropy decodes one after another on the decode stream; the pixel kernel of
 * every run but the last runs on a second stream of the batch as soon as its
 * run's entropy is done, beside the next runs' entropy decodes.  The decode
 * stream waits for the second stream, so the decode is complete when its
 * stream is.  Results are the same for every value.  parts = 0 or 1 runs the
 * two kernels one after the other (the per-kernel timings of
 * cjd_batch_kernel_timings are then each kernel's own); at most 4.  Takes
 * effect at once for the staged batch. */
int cjd_batch_set_overlap(cjd_batch *b, int parts);

/* Per-image tuning by scan size and density (no counterpart in the
 * reference): an image whose entropy-coded scan has at most max_scan_bytes
 * bytes and at most max_bits_per_block scan bits per 8x8 block (0 = any)
 * uses the matching entry with the smallest max_scan_bytes, then the
 * smallest max_bits_per_block: its tile length, checkpoint spacing
 * (0 = default) and tiles per work item.  Images past every entry use the
 * batch options.  n = 0 clears the table.  Entries come from an autotuner
 * (cudajpeg_amd.autotune).  Takes effect at the next cjd_batch_stage. */
typedef struct {
  uint64_t max_scan_bytes;
  int tile_bits, ck_bits, item_tiles, max_bits_per_block;
} cjd_tune_entry;
int cjd_batch_set_tune_table(cjd_batch *b, int n, const cjd_tune_entry *entries);

/* Parse, unstuff (src/decodeScanGPU.cu:11-24) and upload n files; plans the
 * entropy and pixel work lists.  errors[i] gets each file's parse result; a
 * failed file is skipped by the decode.  Returns the first device error. */
int cjd_batch_stage(cjd_batch *b, int n, const uint8_t *const *files, const size_t *sizes, int *errors);

/* cjd_batch_stage with the uploads (and the clear of the batch's scratch)
 * ordered on `stream` (a hipStream_t) and not waited for: staging one batch
 * overlaps another batch's decode on another stream.  The caller orders the
 * batch's next decode after this stream.  No counterpart in the reference
 * (its uploads are synchronous copies, src/decodeScanGPU.cu:146-151). */
int cjd_batch_stage_async(cjd_batch *b, int n, const uint8_t *const *files, const size_t *sizes, int *errors,
                          void *stream);

/* Enqueue the decode of the staged batch on `stream` (asynchronous):
 * entropy decode (replaces decodeScanGPU, src/decodeScanGPU.cu:123) and the
 * fused pixel stage (replaces iDCT_resample_colourTransform,
 * src/pixelTransformGPU.cu:621).  The entropy and pixel launches alone (in
 * parts, see cjd_batch_set_overlap): the batch's scratch (work-item ticket
 * counter, per-image status, lookback records) is not cleared between
 * decodes, since every entropy launch tags what it writes there with an
 * epoch of its own.  Decodes of one batch are ordered by the
 * caller (one stream, or events); different batches are independent. */
int cjd_batch_decode(cjd_batch *b, void *stream);

/* The two halves of cjd_batch_decode.  decode_scan replaces decodeScanGPU
 * (src/decodeScanGPU.cu:123) and waits for it (running the serial decoder for
 * images whose tile links did not resolve); transform replaces
 * iDCT_resample_colourTransform (src/pixelTransformGPU.cu:621) and is
 * asynchronous. */
int cjd_batch_decode_scan(cjd_batch *b, void *stream);
int cjd_batch_transform(cjd_batch *b, void *stream);

/* Number of images decoded again since the batch was created: images the
 * tiled decoder could not resolve or reported an error for are decoded again
 * by the tiled decoder alone, then by the tiled decoder in ordered launches
 * (no work item waits for another of its launch) and, if that still reports
 * anything, by the serial decoder (cjd_batch_status tells which). */
int cjd_batch_fallbacks(cjd_batch *b);

/* Limit of one wait of the tiled decoder's lookback for a predecessor work
 * item, in microseconds (default 200000; 0 = give up unless the predecessor
 * has already published).  A work item whose wait runs out marks its image
 * unresolved and its successors give up in turn; cjd_batch_finish then
 * decodes that image again with the serial decoder, so results do not depend
 * on the limit, and every entropy launch ends.  No counterpart in the
 * reference, whose scan decode is serial (src/decodeScanGPU.cu:435-451). */
int cjd_batch_set_wait_limit(cjd_batch *b, int64_t microseconds);

/* Number of checks after a decode (cjd_batch_finish, cjd_batch_decode_scan)
 * that found a work-item counter off -- some work item was never decoded --
 * and so decoded every image of the batch again with the serial decoder.
 * Zero in a correct run.  No counterpart in the reference. */
int cjd_batch_protocol_redos(cjd_batch *b);

/* Status bits of image i after the last cjd_batch_finish / decode_scan: 1 an
 * invalid code or coefficient index (SYNTAX_ERROR), 2 the scan ended before
 * the frame's blocks, 4 the tiled decoder could not resolve the image, 8 the
 * image's entries did not fit the tiled decoder's streams; after a redo, 0x200
 * the tiled decoder decoded the image again alone, 0x400 it did so in ordered
 * launches, or 0x100 the serial decoder did (bits 0-7 are then that
 * decode's), and bits 16-23 hold the first launch's bits that led to the
 * redo.  For diagnostics; no counterpart in the
 * reference. */
int cjd_batch_status(cjd_batch *b, int i, uint32_t *bits);

/* Wait for the decode, run the serial entropy decoder for any image whose
 * tile links did not resolve, and report per-image reference error codes. */
int cjd_batch_finish(cjd_batch *b, void *stream, int *errors);

/* The batch's decode captured as one HIP graph on `stream` (a created
 * stream, not the null stream; nothing runs), and replays of it on any
 * stream.  The graph's first kernel clears the batch's launch counters,
 * status words and lookback states, so every replay's entropy launches take
 * ticket base 0 and one fixed epoch.  cjd_batch_finish checks a replay and
 * redoes images exactly as after cjd_batch_decode (direct launches).  A
 * replay returns CJD_UNSUPPORTED_ERROR once the batch no longer matches the
 * graph (re-staged with other sizes or into reallocated buffers, or other
 * options, parts or wait limit): capture again then.  Re-staging files of the
 * same shape (a size bucket) keeps the graph valid.  No counterpart in the
 * reference, which launches its kernels one by one (src/decodeScanGPU.cu:
 * 156-220). */
int cjd_batch_capture(cjd_batch *b, void *stream);
int cjd_batch_replay(cjd_batch *b, void *stream);

/* Device time of the last cjd_batch_decode (ms), measured with HIP events on
 * the decode stream: entropy stage, pixel kernel, whole decode. */
int cjd_batch_timings(cjd_batch *b, float *ms_entropy, float *ms_pixel, float *ms_total);
/* The entropy stage of the last cjd_batch_decode split by kernel (ms): the
 * tiled decode kernel alone, and the rest of the stage after it (0 unless a
 * kernel runs there; the record expansion runs when coefficients are asked
 * for).  Tiled entropy mode only.  No counterpart in the reference; used by
 * the benchmark's per-kernel roofline. */
int cjd_batch_kernel_timings(cjd_batch *b, float *ms_entropy_kernel, float *ms_expand);

/* Geometry and device pointers of image i's outputs.  The pixel kernel reads
 * the tiled decoder's block records, not dense coefficients: the first
 * CJD_OUT_COEFFICIENTS request after a decode fills the coefficient buffer of
 * the whole batch from them (a kernel on the decode's stream, waited for). */
int cjd_batch_image_info(cjd_batch *b, int i, cjd_frame_info *info);
int cjd_batch_device_ptr(cjd_batch *b, int i, int what, void **ptr, size_t *bytes);
/* Copy image i's output `what` to host memory (synchronous). */
int cjd_batch_download(cjd_batch *b, int i, int what, void *host, size_t cap);
/* Totals of the staged batch: images, luma pixels, scan bytes (unstuffed),
 * 8x8 blocks, output bytes. */
int cjd_batch_totals(cjd_batch *b, int64_t *images, int64_t *pixels, int64_t *scan_bytes, int64_t *blocks,
                     int64_t *out_bytes);

/* Allocation-free decode of one image (no counterpart in the reference: for
 * callers that cannot allocate device memory -- tight serving loops, harnesses
 * that forbid hipMalloc).  The same path and kernels as cjd_batch_stage +
 * cjd_batch_decode + cjd_batch_finish for one file (parse, unstuff, tiled
 * entropy decode with the serial redo, fused pixel stage), but every device
 * buffer of the decode lives in one static device array of
 * cjd_noalloc_scratch_bytes() bytes (256 MiB) that each GPU the library runs
 * on holds until the process ends; nothing is allocated on the device, and
 * host-side planner buffers are reused from call to call.  `data`/`size` is
 * the file in host memory; the RGB (W*H*3) or gray (W*H) frame is written to
 * the caller's device buffer `frame` (frame_cap bytes, 16-byte aligned).  Ordered on `stream`
 * and synchronised before returning.  Returns the reference error code;
 * CJD_OOM_ERROR when the decode needs more scratch than the array holds
 * (*scratch_needed, if not NULL, gets the bytes it needs: about 60 x the
 * entropy-coded bytes plus 136 bytes per 8x8 block).  Every call on a device
 * shares that device's array; calls on one device are serialised by the
 * library (each holds the device's array until its decode has finished), and
 * the array's contents after a call are not defined. */
int cjd_decode_noalloc(int device, const uint8_t *data, size_t size, int idct_mode, void *frame, size_t frame_cap,
                       void *stream, size_t *scratch_needed);
size_t cjd_noalloc_scratch_bytes(void);

/* Human-readable name of a reference error code (src/format.cu:62-78). */
const char *cjd_error_name(int code);

/* ---------------------------------------------------------------- serving --
 * A request loop (no counterpart in the reference, which decodes one file per
 * openJPG call, src/main.cu:18-50): JPEG byte streams are submitted as they
 * arrive; a batcher thread flushes when max_batch requests wait or the oldest
 * has waited max_wait_us, admits requests while the decoded frames held by
 * the loop fit frame_budget bytes, stages each batch while earlier ones
 * decode (`slots` batches, one stream each), and a completion thread hands
 * finished frames to cjd_server_poll, batch by batch in launch order.  Frames stay in device memory until
 * cjd_server_release; every completed request (also one with an error) must
 * be released. */
typedef struct cjd_server cjd_server;

typedef struct {
  int max_batch;          /* flush at this many waiting requests (0 = 64) */
  int max_wait_us;        /* ... or when the oldest has waited this long (0 = 2000) */
  uint64_t frame_budget;  /* bytes of frames admitted and not released (0 = 64 GiB) */
  int slots;              /* batches in flight or holding frames (0 = 4) */
  int idct_mode;          /* CJD_IDCT_* */
  int tile_bits;          /* entropy tile length (0 = default) */
  int extensions;         /* CJD_EXT_* the loop accepts (0 = refuse those files as the reference does) */
} cjd_server_options;

typedef struct {
  uint64_t id;              /* from cjd_server_submit */
  int error;                /* reference error code (include/format.h numbering) */
  int batch_size;           /* requests decoded in the same launch */
  void *frame;              /* device pointer to the RGB / gray frame (NULL on error) */
  uint64_t frame_bytes;
  int width, height, channels;
  int64_t t_arrival_ns;     /* host monotonic clock: submit */
  int64_t t_launch_ns;      /*   batch launched */
  int64_t t_ready_ns;       /*   frame ready in device memory */
} cjd_completion;

int cjd_server_create(cjd_context *ctx, const cjd_server_options *opt, cjd_server **out);
/* Copies the bytes; *id identifies the request.  Returns a parse error for a
 * file without a decodable frame header (nothing is queued then). */
int cjd_server_submit(cjd_server *s, const uint8_t *data, size_t size, uint64_t *id);
/* Up to max completions, waiting up to wait_us for the first; returns the count. */
int cjd_server_poll(cjd_server *s, cjd_completion *out, int max, int wait_us);
int cjd_server_release(cjd_server *s, uint64_t id);
/* Counters of the loop since it was created, for monitoring: requests
 * submitted and completed, requests waiting for a batch, batches launched and
 * not yet finished, images the slots' batches decoded again
 * (cjd_batch_fallbacks summed), batches launched, and host time (ns) the
 * batcher spent staging and launching batches, the completion thread spent
 * in cjd_batch_finish, and the batcher waited for a free slot; then the
 * batches' device time (ns, cjd_batch_timings of each decode summed, and its
 * largest value) and the time from a batch's launch to its completion
 * (summed). */
typedef struct {
  int64_t submitted, completed, pending, inflight, redone;
  int64_t batches, stage_ns, launch_ns, finish_ns, slot_wait_ns;
  int64_t device_ns, device_max_ns, turnaround_ns;
  int64_t finish_max_ns;  /* longest finish of one batch (its redos included) */
  int64_t protocol_redos; /* finishes that found a launch's counters off and redid the whole batch */
} cjd_server_counters;
int cjd_server_stats(cjd_server *s, cjd_server_counters *out);
/* Host monotonic clock of the loop's timestamps, in ns. */
int64_t cjd_server_now_ns(void);
/* Replays an arrival trace: file k is submitted at start + arrival_ns[k]
 * (non-decreasing), every frame is released as it completes;
 * latency_ns[k] = ready - submit, errors[k] its code; *wall_ns = last ready -
 * start.  For load tests of the loop. */
int cjd_server_replay(cjd_server *s, int n, const uint8_t *const *files, const size_t *sizes,
                      const int64_t *arrival_ns, int64_t *latency_ns, int *errors, int64_t *wall_ns);
void cjd_server_destroy(cjd_server *s);

#ifdef __cplusplus
}
#endif
#endif
