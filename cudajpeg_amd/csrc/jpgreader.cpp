// Reference-style reader API (include/format.h) over the GPU engine.
//
// openJPG / openJPGMem run the marker loop of src/format.cu:113-215; the
// segment parsers share their checks with the C ABI parser (parse.cpp); at SOS
// decodeScanGPU stages the file on the device and runs the entropy decode; at
// EOI iDCT_resample_colourTransform runs the fused pixel stage and copies the
// RGB (or gray) frame into the reader's host buffers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/cudajpeg_amd.h"
#include "../../include/format.h"
#include "parse_internal.h"

namespace {

struct ReaderEngine {
  cjd_context* ctx = nullptr;
  cjd_batch* batch = nullptr;
  cjd_frame_info f{};
  bool staged = false;
};

ReaderEngine* eng(JPGReader* jpg) { return static_cast<ReaderEngine*>(jpg->engine); }

void set_error(JPGReader* jpg, int code, int line, const char* file, const char* func) {
  jpg->error = code;
  jpg->error_line = line;
  jpg->error_file = file;
  jpg->error_func = func;
}

#define RTHROW(e)                                               \
  do {                                                          \
    set_error(jpg, (e), __LINE__, "jpgreader.cpp", __func__);   \
    return;                                                     \
  } while (0)

// Run one of the shared segment parsers at jpg->pos.
template <class F>
void run_parser(JPGReader* jpg, F fn, const char* func) {
  cjd::ParseState st;
  st.ext = jpg->extensions;
  st.buf = jpg->buf;
  st.pos = jpg->pos;
  st.end = jpg->end;
  st.f = &eng(jpg)->f;
  fn(st);
  if (st.error) {
    set_error(jpg, st.error, st.error_line, "parse.cpp", func);
    return;
  }
  jpg->pos = const_cast<unsigned char*>(st.pos);
}

// Grow-only host buffer of `need` elements (the ensureMemSize policy of
// src/format.cu:81-110): size = need, max_size = what is allocated.
template <class M>
int grow(M* m, size_t need) {
  if (need > 0xFFFFFFFFu) return OOM_ERROR;
  if (!m->mem || m->max_size < need) {
    free(m->mem);
    m->mem = static_cast<decltype(m->mem)>(malloc((need ? need : 1) * sizeof(*m->mem)));
    m->max_size = m->mem ? (unsigned int)need : 0u;
    if (!m->mem) {
      m->size = 0;
      return OOM_ERROR;
    }
  }
  m->size = (unsigned int)need;
  return 0;
}

}  // namespace

extern "C" {

unsigned short read16(const unsigned char* pos) { return (unsigned short)((pos[0] << 8) | pos[1]); }

JPGReader* newJPGReader(void) {
  JPGReader* r = static_cast<JPGReader*>(calloc(1, sizeof(JPGReader)));
  if (!r) return nullptr;
  ReaderEngine* e = new (std::nothrow) ReaderEngine;
  if (!e || cjd_context_create(0, &e->ctx) || cjd_batch_create(e->ctx, &e->batch)) {
    if (e) {
      if (e->batch) cjd_batch_destroy(e->batch);
      if (e->ctx) cjd_context_destroy(e->ctx);
      delete e;
    }
    free(r);
    return nullptr;
  }
  r->engine = e;
  r->idct_mode = CJD_IDCT_INT;
  r->entropy_mode = CJD_ENTROPY_TILED;
  return r;
}

void delJPGReader(JPGReader* r) {
  if (!r) return;
  ReaderEngine* e = eng(r);
  if (e) {
    cjd_batch_destroy(e->batch);
    cjd_context_destroy(e->ctx);
    delete e;
  }
  free(r->pixels);
  free(r->file_buf.mem);
  for (int i = 0; i < 3; ++i) {
    free(r->channels[i].working_space.mem);
    free(r->channels[i].raw_pixels.mem);
    free(r->channels[i].pixels.mem);
  }
  free(r);
}

void printError(JPGReader* r) {
  printf("Code: %s", cjd_error_name(r->error));
  if (r->error_func) printf(" Func: %s", r->error_func);
  if (r->error_file) printf(" File: %s", r->error_file);
  if (r->error_line) printf(" Line: %d", r->error_line);
}

void skipBlock(JPGReader* jpg) { jpg->pos += read16(jpg->pos); }

void decodeSOF(JPGReader* jpg) {
  run_parser(jpg, cjd::parse_sof, __func__);
  if (jpg->error) return;
  const cjd_frame_info& f = eng(jpg)->f;
  jpg->width = (unsigned short)f.width;
  jpg->height = (unsigned short)f.height;
  jpg->num_channels = (unsigned char)f.num_channels;
  jpg->num_blocks_x = (unsigned short)f.mcus_x;
  jpg->num_blocks_y = (unsigned short)f.mcus_y;
  jpg->block_size_x = (unsigned short)f.mcu_width;
  jpg->block_size_y = (unsigned short)f.mcu_height;
  for (int i = 0; i < f.num_channels; ++i) {
    ColourChannel* c = &jpg->channels[i];
    c->id = f.comp_id[i];
    c->dq_id = f.dq_id[i];
    c->samples_x = f.samples_x[i];
    c->samples_y = f.samples_y[i];
    c->width = f.comp_width[i];
    c->height = f.comp_height[i];
    c->stride = f.comp_stride[i];
    c->plane_height = f.comp_plane_height[i];
  }
  // output buffers (src/format.cu:332-345)
  if (f.num_channels == 3) {
    const size_t n = (size_t)f.width * f.height * 3;
    if (!jpg->pixels || jpg->max_pixels_size < (int)n) {
      free(jpg->pixels);
      jpg->pixels = static_cast<unsigned char*>(malloc(n));
      if (!jpg->pixels) RTHROW(OOM_ERROR);
      jpg->max_pixels_size = (int)n;
    }
  } else {
    if (grow(&jpg->channels[0].pixels, (size_t)f.width * f.height)) RTHROW(OOM_ERROR);
  }
}

void decodeDHT(JPGReader* jpg) { run_parser(jpg, cjd::parse_dht, __func__); }

void decodeDQT(JPGReader* jpg) {
  run_parser(jpg, cjd::parse_dqt, __func__);
  if (!jpg->error) memcpy(jpg->dq_tables, eng(jpg)->f.dq_tables, sizeof(jpg->dq_tables));
}

void decodeDRI(JPGReader* jpg) {
  run_parser(jpg, cjd::parse_dri, __func__);
  if (!jpg->error) jpg->restart_interval = eng(jpg)->f.restart_interval;
}

// SOS header, then the whole scan on the GPU with the given entropy decoder;
// working_space is filled when `coefficients`.  The clock() ticks it takes
// are added to jpg->time (src/decodeScanGPU.cu:154, 224-225).
static void decode_scan(JPGReader* jpg, int entropy_mode, bool coefficients) {
  const clock_t t0 = clock();
  ReaderEngine* e = eng(jpg);
  // CJD_EXT_NONINTERLEAVED: a one-component first scan; CJD_EXT_PROGRESSIVE:
  // a progressive frame's first scan.  The batch's planner parses the file's
  // scans.
  const bool one = e->f.progressive || ((jpg->extensions & CJD_EXT_NONINTERLEAVED) && jpg->num_channels > 1 &&
                                        jpg->pos + 3 <= jpg->end && jpg->pos[2] == 1);
  if (!one) {
    run_parser(jpg, cjd::parse_sos, "decodeScan");
    if (jpg->error) return;
    for (int i = 0; i < jpg->num_channels; ++i) {
      jpg->channels[i].dc_id = e->f.dc_id[i];
      jpg->channels[i].ac_id = e->f.ac_id[i];
    }
  }
  int err = cjd_batch_set_options(e->batch, jpg->idct_mode, entropy_mode, 0, jpg->keep_planes);
  if (!err) err = cjd_batch_set_extensions(e->batch, jpg->extensions);
  if (err) RTHROW(err);
  const uint8_t* files[1] = {jpg->buf};
  const size_t sizes[1] = {(size_t)(jpg->end - jpg->buf) + 2};
  int ferr = 0;
  err = cjd_batch_stage(e->batch, 1, files, sizes, &ferr);
  if (err) RTHROW(err);
  if (ferr) RTHROW(ferr);
  err = cjd_batch_decode_scan(e->batch, nullptr);
  if (err) RTHROW(err);
  cjd_frame_info info;
  ferr = cjd_batch_image_info(e->batch, 0, &info);
  if (ferr) RTHROW(ferr);
  e->staged = true;
  if (coefficients) {
    // dequantised working_space in the reference layout (src/decodeScanCPU.cu:78-96):
    // per component, 64 ints per block in decode order, natural order within a block
    std::vector<int16_t> q((size_t)info.num_blocks * 64);
    err = cjd_batch_download(e->batch, 0, CJD_OUT_COEFFICIENTS, q.data(), q.size() * 2);
    if (err) RTHROW(err);
    static const unsigned char zz[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
    int* wpos[3] = {nullptr, nullptr, nullptr};
    for (int c = 0; c < jpg->num_channels; ++c) {
      ColourChannel* ch = &jpg->channels[c];
      if (grow(&ch->working_space, (size_t)ch->stride * ch->plane_height)) RTHROW(OOM_ERROR);
      wpos[c] = ch->working_space.mem;
    }
    // the staged file's tables (one-component scans: each component's as
    // latched at its scan)
    size_t blk = 0;
    for (int m = 0; m < info.mcus_x * info.mcus_y; ++m)
      for (int c = 0; c < jpg->num_channels; ++c) {
        const uint16_t* dq = info.dq_tables16[info.dq_id[c]];
        for (int k = 0; k < jpg->channels[c].samples_x * jpg->channels[c].samples_y; ++k, ++blk) {
          const int16_t* src = &q[blk * 64];
          for (int z = 0; z < 64; ++z) wpos[c][zz[z]] = (int)src[zz[z]] * dq[z];
          wpos[c] += 64;
        }
      }
  }
  jpg->pos = jpg->end;
  jpg->time += clock() - t0;
}

// include/decodeScanGPU.h:6 -- the scan with the reader's entropy decoder
// (the tiled decoder by default).
void decodeScanGPU(JPGReader* jpg) { decode_scan(jpg, jpg->entropy_mode, jpg->keep_coefficients != 0); }

// include/decodeScanCPU.h:6 -- the reference's serial scan decoder, whose
// result is working_space (src/decodeScanCPU.cu:98-152): here the serial
// decoder on the device (one lane per restart segment, the reference's
// order), and working_space is always filled.
void decodeScanCPU(JPGReader* jpg) { decode_scan(jpg, CJD_ENTROPY_SERIAL, true); }

// include/pixelTransformGPU.h:17 -- fused dequant + iDCT + upsample + colour.
void iDCT_resample_colourTransform(JPGReader* jpg) {
  ReaderEngine* e = eng(jpg);
  if (!e->staged) RTHROW(SYNTAX_ERROR);  // EOI without a scan
  int err = cjd_batch_transform(e->batch, nullptr);
  if (err) RTHROW(err);
  if (hipStreamSynchronize(nullptr) != hipSuccess) RTHROW(CUDA_KERNEL_LAUNCH_ERROR);
  const size_t n = (size_t)jpg->width * jpg->height * (jpg->num_channels == 3 ? 3 : 1);
  unsigned char* dst = jpg->num_channels == 3 ? jpg->pixels : jpg->channels[0].pixels.mem;
  err = cjd_batch_download(e->batch, 0, CJD_OUT_PIXELS, dst, n);
  if (err) RTHROW(err);
  if (jpg->keep_planes) {
    for (int c = 0; c < jpg->num_channels; ++c) {
      ColourChannel* ch = &jpg->channels[c];
      const size_t pn = (size_t)ch->stride * ch->plane_height;
      if (grow(&ch->raw_pixels, pn)) RTHROW(OOM_ERROR);
      err = cjd_batch_download(e->batch, 0, CJD_OUT_PLANE0 + c, ch->raw_pixels.mem, pn);
      if (err) RTHROW(err);
    }
  }
}

// src/format.cu:160-212 on a file already in jpg->file_buf.
static int run_marker_loop(JPGReader* jpg, size_t size) {
  unsigned char* buf = jpg->file_buf.mem;
  jpg->buf = buf;
  jpg->time = 0;
  jpg->error = NO_ERROR;
  jpg->error_line = 0;
  jpg->error_file = jpg->error_func = nullptr;
  if (size < 4 || buf[0] != 0xFF || buf[1] != 0xD8 || buf[size - 2] != 0xFF || buf[size - 1] != 0xD9)
    return FILE_ERROR;
  if (size < 6) return SYNTAX_ERROR;
  jpg->end = buf + size - 2;
  jpg->pos = buf + 2;
  jpg->restart_interval = 0;
  jpg->num_blocks_x = jpg->num_blocks_y = 0;
  ReaderEngine* e = eng(jpg);
  memset(&e->f, 0, sizeof(e->f));
  e->staged = false;
  while (!jpg->error) {
    if (jpg->pos > jpg->end || jpg->pos[0] != 0xFF) {
      set_error(jpg, SYNTAX_ERROR, __LINE__, "jpgreader.cpp", "openJPG");
      break;
    }
    jpg->pos += 2;
    switch (jpg->pos[-1]) {
      case 0xC0: decodeSOF(jpg); break;
      case 0xC2:  // CJD_EXT_PROGRESSIVE; src/format.cu:198-200 rejects it as an unknown marker
        if (!(jpg->extensions & CJD_EXT_PROGRESSIVE)) {
          set_error(jpg, SYNTAX_ERROR, __LINE__, "jpgreader.cpp", "openJPG");
          break;
        }
        decodeSOF(jpg);
        e->f.progressive = 1;
        break;
      case 0xC4: decodeDHT(jpg); break;
      case 0xDB: decodeDQT(jpg); break;
      case 0xDD: decodeDRI(jpg); break;
      case 0xDA: decodeScanGPU(jpg); break;
      case 0xFE: skipBlock(jpg); break;
      case 0xD9: break;
      default:
        if ((jpg->pos[-1] & 0xF0) == 0xE0) skipBlock(jpg);
        else set_error(jpg, SYNTAX_ERROR, __LINE__, "jpgreader.cpp", "openJPG");
    }
    if (jpg->error) break;
    if (jpg->pos[-1] == 0xD9) {
      iDCT_resample_colourTransform(jpg);
      break;
    }
  }
  return jpg->error;
}

int openJPGMem(JPGReader* jpg, const unsigned char* data, size_t size) {
  if (!jpg || !data) return FILE_ERROR;
  if (grow(&jpg->file_buf, size + 16)) return OOM_ERROR;
  memcpy(jpg->file_buf.mem, data, size);
  return run_marker_loop(jpg, size);
}

int openJPG(JPGReader* jpg, const char* filename) {
  FILE* f = fopen(filename, "rb");
  if (!f) return FILE_ERROR;
  fseek(f, 0, SEEK_END);
  const long size = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (size <= 0 || grow(&jpg->file_buf, (size_t)size + 16)) {
    fclose(f);
    return size <= 0 ? FILE_ERROR : OOM_ERROR;
  }
  const size_t got = fread(jpg->file_buf.mem, 1, (size_t)size, f);
  fclose(f);
  if (got != (size_t)size) return FILE_ERROR;
  return run_marker_loop(jpg, (size_t)size);
}

// src/format.cu:218-229
void writeJPG(JPGReader* jpg, const char* filename) {
  FILE* f = fopen(filename, "wb");
  if (!f) {
    printf("Couldn't open output file %s\n", filename);
    return;
  }
  fprintf(f, "P%d\n%d %d\n255\n", jpg->num_channels > 1 ? 6 : 5, jpg->width, jpg->height);
  fwrite(jpg->num_channels == 1 ? jpg->channels[0].pixels.mem : jpg->pixels, 1,
         (size_t)jpg->width * jpg->height * jpg->num_channels, f);
  fclose(f);
}

}  // extern "C"
