/*
 * Host API of the MI355X baseline-JPEG decoder, keeping the reference's
 * reader interface (include/format.h:1-123 of jndean/CUDAjpeg): JPGReader,
 * openJPG, the marker / quant-table / Huffman-table parsers, the decodeScan
 * entry points and the planar/RGB output buffers.  The scan decode and the
 * pixel stage run on the GPU through the C ABI of cudajpeg_amd.h.
 *
 * Differences from the reference, by design:
 *  - decodeDHT keeps BITS/HUFFVAL (compact device tables are built from them)
 *    instead of filling 65536-entry LUTs (src/format.cu:371-391);
 *  - coefficients stay on the device; `keep_coefficients` copies the
 *    dequantised working_space back in the reference layout for inspection;
 *  - openJPGMem decodes a file already in memory;
 *  - decodeScanCPU, the reference's serial scan decoder, runs the serial
 *    decoder on the device (one lane per restart segment) and fills
 *    working_space; there is no host decode path.
 */
#ifndef CUDAJPEG_AMD_FORMAT_H
#define CUDAJPEG_AMD_FORMAT_H
#include <stddef.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/format.h:7-14 */
#define NO_ERROR 0
#define SYNTAX_ERROR 1
#define UNSUPPORTED_ERROR 2
#define OOM_ERROR 3
#define CUDA_MEM_ERROR 4
#define FILE_ERROR 5
#define CUDA_KERNEL_LAUNCH_ERROR 6
#define PROGRAMMER_ERROR -1

/* include/format.h:36-52: a host buffer, `size` elements in use of
 * `max_size` allocated (grow-only). */
typedef struct _ManagedUCharMem {
  unsigned char *mem;
  unsigned int size, max_size;
} ManagedUCharMem;

typedef struct _ManagedIntMem {
  int *mem;
  unsigned int size, max_size;
} ManagedIntMem;

/* include/format.h:55-69 */
typedef struct _ColourChannel {
  int id;
  int dq_id, ac_id, dc_id;
  int width, height;
  int samples_x, samples_y, stride;
  int plane_height;
  ManagedIntMem working_space;     /* host dequantised coefficients (keep_coefficients, decodeScanCPU) */
  ManagedUCharMem raw_pixels;      /* host planar samples (keep_planes), stride x plane_height */
  ManagedUCharMem pixels;          /* grayscale output, width x height */
} ColourChannel;

/* include/format.h:72-103 */
typedef struct _JPGReader {
  unsigned char *buf, *pos, *end;  /* file buffer (owned by file_buf) */
  unsigned short width, height;
  unsigned short num_blocks_x, num_blocks_y;   /* MCUs */
  unsigned short block_size_x, block_size_y;   /* MCU size in pixels */
  unsigned char num_channels;
  int restart_interval;
  ColourChannel channels[3];
  unsigned char *pixels;           /* RGB output, width x height x 3 */
  int max_pixels_size;
  unsigned char dq_tables[4][64];
  ManagedUCharMem file_buf;
  /* options */
  int idct_mode;                   /* CJD_IDCT_INT / _FLOAT / _MFMA */
  int entropy_mode;                /* CJD_ENTROPY_TILED / _SERIAL */
  int keep_coefficients;
  int keep_planes;
  int extensions;                  /* CJD_EXT_* accepted beyond the reference (0 = none) */
  void *engine;                    /* device context, batch and parsed frame */
  /* clock() ticks of the last openJPG's scan decode (include/format.h:100;
   * set to 0 by openJPG, src/format.cu:174, and accumulated by the scan
   * decode, src/decodeScanGPU.cu:154, 224-225) */
  clock_t time;
  /* error handling (include/format.h:22-27, :100-102) */
  int error, error_line;
  const char *error_file, *error_func;
} JPGReader;

JPGReader *newJPGReader(void);                                /* src/format.cu:12 */
void delJPGReader(JPGReader *reader);                          /* src/format.cu:32 */
int openJPG(JPGReader *reader, const char *filename);          /* src/format.cu:113 */
int openJPGMem(JPGReader *reader, const unsigned char *data, size_t size);
void writeJPG(JPGReader *jpg, const char *filename);           /* src/format.cu:218 */
void printError(JPGReader *reader);                            /* src/format.cu:62 */
unsigned short read16(const unsigned char *pos);               /* src/format.cu:232 */
void skipBlock(JPGReader *jpg);                                /* src/format.cu:237 */
void decodeSOF(JPGReader *jpg);                                /* src/format.cu:242 */
void decodeDHT(JPGReader *jpg);                                /* src/format.cu:357 */
void decodeDQT(JPGReader *jpg);                                /* src/format.cu:424 */
void decodeDRI(JPGReader *jpg);                                /* src/format.cu:404 */
void decodeScanGPU(JPGReader *jpg);                            /* include/decodeScanGPU.h:6 */
void decodeScanCPU(JPGReader *jpg);                            /* include/decodeScanCPU.h:6 */
void iDCT_resample_colourTransform(JPGReader *jpg);            /* include/pixelTransformGPU.h:17 */

#ifdef __cplusplus
}
#endif
#endif
