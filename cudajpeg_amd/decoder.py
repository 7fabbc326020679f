"""Python face of the decoder: a batch decoder over the C ABI and the
reference-style JPGReader (include/format.h).

Everything runs on the GPU through libcudajpeg_amd.so; this module only moves
arguments and copies results out.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def _stream_handle(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(getattr(stream, "cuda_stream"))


class DecodeError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        super().__init__(f"{what}{L.error_name(code)} ({code})")
        self.code = code


def _check(code: int, what: str = "") -> None:
    if code:
        raise DecodeError(code, what)


def parse(data: bytes, ext: int = 0) -> tuple[int, L.FrameInfo]:
    """Host marker/table parse of one file (cjd_parse, or cjd_parse_ext with
    the extensions `ext`)."""
    info = L.FrameInfo()
    e = L.lib().cjd_parse_ext(data, len(data), ext, ctypes.byref(info))
    return e, info


class Context:
    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        self._destroy = L.lib().cjd_context_destroy  # see BatchDecoder.__init__
        _check(L.lib().cjd_context_create(device, ctypes.byref(self.h)), "context: ")

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class ImageOut:
    width: int
    height: int
    channels: int
    num_blocks: int
    error: int


class BatchDecoder:
    """Decodes a batch of JPEG byte strings together on one GPU.

    stage() parses, unstuffs and uploads; decode() enqueues the entropy and
    pixel kernels; finish() waits and reports per-image reference error codes.
    """

    def __init__(self, idct: int = L.IDCT_INT, entropy: int = L.ENTROPY_TILED, tile_bits: int = 0,
                 planes: bool = False, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.h = ctypes.c_void_p()
        # kept here: at interpreter exit the module's globals may be gone
        # before this object's finaliser runs
        self._destroy = L.lib().cjd_batch_destroy
        _check(L.lib().cjd_batch_create(self.ctx.h, ctypes.byref(self.h)), "batch: ")
        self.set_options(idct, entropy, tile_bits, planes)
        self.n = 0
        self.errors = np.zeros(0, np.int32)

    def set_options(self, idct: int, entropy: int = L.ENTROPY_TILED, tile_bits: int = 0, planes: bool = False):
        _check(L.lib().cjd_batch_set_options(self.h, idct, entropy, tile_bits, int(planes)), "options: ")

    def set_tuning(self, ck_bits: int = 0, item_tiles: int = 0):
        """Tiled entropy decoder: checkpoint spacing in bits and tiles per
        work item; 0 = defaults; applies from the next stage()."""
        _check(L.lib().cjd_batch_set_tuning(self.h, ck_bits, item_tiles), "tuning: ")

    def set_extensions(self, ext: int) -> None:
        """Extensions (EXT_*) accepted at the next stage; 0 = the reference's parsing."""
        _check(L.lib().cjd_batch_set_extensions(self.h, ext), "extensions: ")

    def set_unstuff(self, mode: int) -> None:
        """UNSTUFF_DEVICE (default): scans unstuffed by a kernel at staging;
        UNSTUFF_HOST: by the host planner.  From the next stage."""
        _check(L.lib().cjd_batch_set_unstuff(self.h, mode), "unstuff: ")

    def set_overlap(self, parts: int = 2) -> None:
        """Decode in parts (default 2; 1 = off): each run's pixel kernel runs
        beside the next runs' entropy decodes, on a second stream."""
        _check(L.lib().cjd_batch_set_overlap(self.h, int(parts)), "overlap: ")

    def set_tune_table(self, table) -> None:
        """Per-image entropy parameters by scan size and density: a list of
        (max_scan_bytes, tile_bits, ck_bits, item_tiles[, max_bits_per_block]),
        max_bits_per_block 0 or absent = any density."""
        arr = (L.TuneEntry * max(1, len(table)))(*[L.TuneEntry(*t[:4], int(t[4]) if len(t) > 4 else 0)
                                                    for t in table])
        _check(L.lib().cjd_batch_set_tune_table(self.h, len(table), arr), "tune table: ")

    def stage(self, files: list[bytes], stream=None, asynchronous: bool = False) -> np.ndarray:
        """Parse, unstuff and plan the files and upload them (the files are
        read during the call, not after).  asynchronous: the uploads are
        ordered on `stream` and not waited for (cjd_batch_stage_async)."""
        n = len(files)
        keep = [ctypes.c_char_p(bytes(f)) for f in files]  # points at each bytes object's own buffer
        ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(k, ctypes.c_void_p) for k in keep])
        sizes = (ctypes.c_size_t * n)(*[len(f) for f in files])
        self.errors = np.zeros(n, np.int32)
        if asynchronous:
            _check(L.lib().cjd_batch_stage_async(self.h, n, ptrs, sizes, self.errors.ctypes.data,
                                                 _stream_handle(stream)), "stage: ")
        else:
            _check(L.lib().cjd_batch_stage(self.h, n, ptrs, sizes, self.errors.ctypes.data), "stage: ")
        del keep
        self.n = n
        return self.errors.copy()

    def decode(self, stream=None) -> None:
        _check(L.lib().cjd_batch_decode(self.h, _stream_handle(stream)), "decode: ")

    def decode_scan(self, stream=None) -> None:
        _check(L.lib().cjd_batch_decode_scan(self.h, _stream_handle(stream)), "decode_scan: ")

    def transform(self, stream=None) -> None:
        _check(L.lib().cjd_batch_transform(self.h, _stream_handle(stream)), "transform: ")

    def finish(self, stream=None) -> np.ndarray:
        errs = np.zeros(max(self.n, 1), np.int32)
        _check(L.lib().cjd_batch_finish(self.h, _stream_handle(stream), errs.ctypes.data), "finish: ")
        self.errors = errs[: self.n]
        return self.errors.copy()

    def capture(self, stream) -> None:
        _check(L.lib().cjd_batch_capture(self.h, _stream_handle(stream)), "capture: ")

    def replay(self, stream) -> None:
        _check(L.lib().cjd_batch_replay(self.h, _stream_handle(stream)), "replay: ")

    def fallbacks(self) -> int:
        return int(L.lib().cjd_batch_fallbacks(self.h))

    def set_wait_limit(self, microseconds: int) -> None:
        """Limit of one lookback wait of the tiled decoder (0 = give up unless
        the predecessor has published; results do not depend on it)."""
        _check(L.lib().cjd_batch_set_wait_limit(self.h, int(microseconds)), "set_wait_limit")

    def status(self, i: int) -> int:
        """Status bits of image i after the last finish (1 syntax, 2 truncated,
        4 unresolved, 8 entries did not fit; after a redo 0x200 decoded again
        by the tiled decoder alone, 0x400 in ordered launches, 0x100 by the
        serial decoder, and bits 16-23 the first launch's bits)."""
        v = ctypes.c_uint32()
        _check(L.lib().cjd_batch_status(self.h, i, ctypes.byref(v)), "status")
        return int(v.value)

    def protocol_redos(self) -> int:
        """Checks after a decode that found a work item never decoded (the
        whole batch was then decoded again serially); 0 in a correct run."""
        return int(L.lib().cjd_batch_protocol_redos(self.h))

    def timings(self) -> tuple[float, float, float]:
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _check(L.lib().cjd_batch_timings(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "timings: ")
        return a.value, b.value, c.value

    def kernel_timings(self) -> tuple[float, float]:
        """(tiled entropy kernel alone, record expansion) of the last decode, ms."""
        a, x = ctypes.c_float(), ctypes.c_float()
        _check(L.lib().cjd_batch_kernel_timings(self.h, ctypes.byref(a), ctypes.byref(x)), "kernel timings: ")
        return a.value, x.value

    def info(self, i: int) -> L.FrameInfo:
        f = L.FrameInfo()
        L.lib().cjd_batch_image_info(self.h, i, ctypes.byref(f))
        return f

    def totals(self) -> dict:
        v = [ctypes.c_int64() for _ in range(5)]
        _check(L.lib().cjd_batch_totals(self.h, *[ctypes.byref(x) for x in v]), "totals: ")
        return dict(zip(["images", "pixels", "scan_bytes", "blocks", "out_bytes"], [x.value for x in v]))

    def device_ptr(self, i: int, what: int = L.OUT_PIXELS) -> tuple[int, int]:
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        _check(L.lib().cjd_batch_device_ptr(self.h, i, what, ctypes.byref(p), ctypes.byref(n)), "device_ptr: ")
        return p.value, n.value

    def _download(self, i: int, what: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        _check(L.lib().cjd_batch_download(self.h, i, what, out.ctypes.data, out.nbytes), "download: ")
        return out

    def pixels(self, i: int) -> np.ndarray:
        f = self.info(i)
        c = 3 if f.num_channels == 3 else 1
        return self._download(i, L.OUT_PIXELS, (f.height, f.width, c), np.uint8)

    def coefficients(self, i: int) -> np.ndarray:
        f = self.info(i)
        return self._download(i, L.OUT_COEFFICIENTS, (f.num_blocks, 64), np.int16)

    def plane(self, i: int, comp: int) -> np.ndarray:
        f = self.info(i)
        return self._download(i, L.OUT_PLANE0 + comp, (f.comp_plane_height[comp], f.comp_stride[comp]), np.uint8)

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def decode(data: bytes, idct: int = L.IDCT_INT) -> np.ndarray:
    """Decode one JPEG to an HxWx{1,3} uint8 array (raises DecodeError)."""
    b = BatchDecoder(idct)
    err = b.stage([data])
    _check(int(err[0]), "parse: ")
    b.decode()
    err = b.finish()
    _check(int(err[0]), "decode: ")
    out = b.pixels(0)
    b.close()
    return out


# ------------------------------------------------ reference-style reader --
class _ManagedUCharMem(ctypes.Structure):
    _fields_ = [("mem", ctypes.POINTER(ctypes.c_uint8)), ("size", ctypes.c_uint), ("max_size", ctypes.c_uint)]


class _ManagedIntMem(ctypes.Structure):
    _fields_ = [("mem", ctypes.POINTER(ctypes.c_int)), ("size", ctypes.c_uint), ("max_size", ctypes.c_uint)]


class _ColourChannel(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int), ("dq_id", ctypes.c_int), ("ac_id", ctypes.c_int), ("dc_id", ctypes.c_int),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("samples_x", ctypes.c_int),
                ("samples_y", ctypes.c_int), ("stride", ctypes.c_int), ("plane_height", ctypes.c_int),
                ("working_space", _ManagedIntMem), ("raw_pixels", _ManagedUCharMem), ("pixels", _ManagedUCharMem)]


class _JPGReader(ctypes.Structure):
    _fields_ = [("buf", ctypes.c_void_p), ("pos", ctypes.c_void_p), ("end", ctypes.c_void_p),
                ("width", ctypes.c_ushort), ("height", ctypes.c_ushort),
                ("num_blocks_x", ctypes.c_ushort), ("num_blocks_y", ctypes.c_ushort),
                ("block_size_x", ctypes.c_ushort), ("block_size_y", ctypes.c_ushort),
                ("num_channels", ctypes.c_ubyte), ("restart_interval", ctypes.c_int),
                ("channels", _ColourChannel * 3),
                ("pixels", ctypes.POINTER(ctypes.c_uint8)), ("max_pixels_size", ctypes.c_int),
                ("dq_tables", (ctypes.c_uint8 * 64) * 4),
                ("file_buf", _ManagedUCharMem),
                ("idct_mode", ctypes.c_int), ("entropy_mode", ctypes.c_int),
                ("keep_coefficients", ctypes.c_int), ("keep_planes", ctypes.c_int), ("extensions", ctypes.c_int),
                ("engine", ctypes.c_void_p), ("time", ctypes.c_long),
                ("error", ctypes.c_int), ("error_line", ctypes.c_int),
                ("error_file", ctypes.c_char_p), ("error_func", ctypes.c_char_p)]


class JPGReader:
    """The reference's reader (newJPGReader / openJPG / delJPGReader,
    src/format.cu:12-215): reusable across images; after open() the RGB or
    gray frame is in `pixels`."""

    def __init__(self, idct: int = L.IDCT_INT, entropy: int = L.ENTROPY_TILED, keep_coefficients: bool = False,
                 keep_planes: bool = False, extensions: int = 0):
        p = L.lib().newJPGReader()
        if not p:
            raise DecodeError(L.OOM_ERROR, "newJPGReader: ")
        self._p = ctypes.cast(p, ctypes.POINTER(_JPGReader))
        r = self._p.contents
        r.idct_mode = idct
        r.entropy_mode = entropy
        r.keep_coefficients = int(keep_coefficients)
        r.keep_planes = int(keep_planes)
        r.extensions = int(extensions)

    @property
    def raw(self) -> _JPGReader:
        return self._p.contents

    def open(self, src) -> int:
        """openJPG(path) or openJPGMem(bytes); returns the reference error code."""
        if isinstance(src, (bytes, bytearray)):
            return int(L.lib().openJPGMem(self._p, bytes(src), len(src)))
        return int(L.lib().openJPG(self._p, str(src).encode()))

    @property
    def error(self) -> tuple[int, int, str, str]:
        r = self.raw
        return (r.error, r.error_line, (r.error_file or b"").decode(), (r.error_func or b"").decode())

    @property
    def pixels(self) -> np.ndarray:
        r = self.raw
        if r.num_channels == 3:
            return np.ctypeslib.as_array(r.pixels, (r.height, r.width, 3)).copy()
        return np.ctypeslib.as_array(r.channels[0].pixels.mem, (r.height, r.width, 1)).copy()

    def working_space(self, comp: int) -> np.ndarray:
        ch = self.raw.channels[comp]
        return np.ctypeslib.as_array(ch.working_space.mem, (ch.working_space.size,)).copy()

    def raw_pixels(self, comp: int) -> np.ndarray:
        ch = self.raw.channels[comp]
        return np.ctypeslib.as_array(ch.raw_pixels.mem, (ch.plane_height, ch.stride)).copy()

    @property
    def time(self) -> int:
        """clock() ticks of the last open's scan decode."""
        return int(self.raw.time)

    def write(self, path: str) -> None:
        L.lib().writeJPG(self._p, path.encode())

    def close(self):
        if getattr(self, "_p", None):
            L.lib().delJPGReader(self._p)
            self._p = None

    def __del__(self):
        self.close()
