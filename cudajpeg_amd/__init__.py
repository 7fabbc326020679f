"""MI355X-native baseline-JPEG decoder (gfx950).

Host parsing in C++, scan decode and the fused pixel stage in hand-written HIP
kernels, behind the C ABI of include/cudajpeg_amd.h and the reference-style
reader API of include/format.h.
"""
from ._lib import (ENTROPY_SERIAL, ENTROPY_TILED, IDCT_FLOAT, IDCT_INT, IDCT_MFMA, OUT_COEFFICIENTS,
                   OUT_PIXELS, OUT_PLANE0, error_name)
from .decoder import BatchDecoder, Context, DecodeError, JPGReader, decode, parse
from .serve import Server, device_frame

__all__ = ["BatchDecoder", "Context", "DecodeError", "JPGReader", "Server", "decode", "device_frame", "parse",
           "error_name",
           "IDCT_INT", "IDCT_FLOAT", "IDCT_MFMA", "ENTROPY_TILED", "ENTROPY_SERIAL",
           "OUT_PIXELS", "OUT_COEFFICIENTS", "OUT_PLANE0"]
