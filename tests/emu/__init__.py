"""ctypes wrapper of the CPU emulator of the decoder's HIP kernels (tests only)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(os.path.join(HERE, "libcjd_emu.so"))
        L.emu_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p]
        L.emu_run.restype = ctypes.c_void_p
        L.emu_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        L.emu_get.restype = ctypes.c_long
        L.emu_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.emu_free.argtypes = [ctypes.c_void_p]
        L.emu_ext.argtypes = [ctypes.c_int]
        L.emu_unstuff_files.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.emu_unstuff_files.restype = ctypes.c_long
        L.emu_unstuff_bytes.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_size_t]
        L.emu_unstuff_bytes.restype = ctypes.c_long
        _lib = L
    return _lib


class EmuResult:
    def __init__(self, files, idct_mode=0, tile_bits=4096, want_planes=False, order=0, serial=False,
                 item_tiles=0, ck_bits=0):
        L = lib()
        n = len(files)
        self._keep = [ctypes.create_string_buffer(f, len(f)) for f in files]
        ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(b, ctypes.c_void_p) for b in self._keep])
        sizes = (ctypes.c_size_t * n)(*[len(f) for f in files])
        self.errors = np.zeros(n, np.int32)
        self.h = L.emu_run(n, ptrs, sizes, idct_mode, tile_bits, int(want_planes), order, int(serial),
                           item_tiles, ck_bits, self.errors.ctypes.data)
        st = np.zeros(4, np.int64)
        L.emu_stats(self.h, st.ctypes.data)
        self.items, self.fixes, self.fallbacks, self.segments = (int(x) for x in st)

    def get(self, i, what, shape, dtype):
        out = np.zeros(shape, dtype)
        n = lib().emu_get(self.h, i, what, out.ctypes.data, out.nbytes)
        if n < 0:
            return None
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().emu_free(self.h)
            self.h = None


def unstuff_files(files):
    """(host words, emulated device words) of the batch, or the error code."""
    n = len(files)
    keep = [ctypes.c_char_p(bytes(f)) for f in files]
    ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(k, ctypes.c_void_p) for k in keep])
    sizes = (ctypes.c_size_t * n)(*[len(f) for f in files])
    cap = sum(len(f) for f in files) + 4096 * n + 200000  # words: data / 4 plus 100 per restart segment, at most
    h = np.zeros(cap, np.uint32)
    d = np.zeros(cap, np.uint32)
    errors = np.zeros(n, np.int32)
    r = lib().emu_unstuff_files(n, ptrs, sizes, h.ctypes.data, d.ctypes.data, cap, errors.ctypes.data)
    if r < 0:
        return r
    return h[:r], d[:r]


def unstuff_bytes(data: bytes, split: bool):
    """(host words, emulated device words) of one raw scan, or the error code."""
    cap = len(data) // 4 + 100 * (len(data) // 2 + 2) + 64  # at most one segment per RSTn pair
    h = np.zeros(cap, np.uint32)
    d = np.zeros(cap, np.uint32)
    r = lib().emu_unstuff_bytes(data, len(data), int(split), h.ctypes.data, d.ctypes.data, cap)
    if r < 0:
        return r
    return h[:r], d[:r]
