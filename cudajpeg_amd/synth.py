"""Seeded synthetic JPEG workloads.

Thin ctypes layer over ``csrc/synth.c`` (the project's own generator and
baseline encoder) plus the workload definitions the decoder is measured on:
the single-image configurations and the mixed-batch / serving distributions.
Every image is a pure function of its parameters and seed.
"""
from __future__ import annotations

import ctypes
import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Params(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("ncomp", ctypes.c_int32),
        ("h_samp", ctypes.c_int32),
        ("v_samp", ctypes.c_int32),
        ("quality", ctypes.c_int32),
        ("restart_interval", ctypes.c_int32),
        ("sigma", ctypes.c_double),
        ("amp_total", ctypes.c_double),
        ("random_tile_frac", ctypes.c_double),
        ("qtables", ctypes.c_int32),
        ("huff", ctypes.c_int32),
        ("dqt16", ctypes.c_int32),
        ("noninterleaved", ctypes.c_int32),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsynth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.POINTER(_Params)
        lib.synth_pixels.argtypes = [P, ctypes.c_void_p]
        lib.synth_pixels.restype = ctypes.c_int
        lib.synth_encode.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                     ctypes.c_void_p, ctypes.c_void_p]
        lib.synth_encode.restype = ctypes.c_long
        lib.synth_num_blocks.argtypes = [P]
        lib.synth_num_blocks.restype = ctypes.c_long
        lib.synth_quant_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        lib.synth_quant_table.restype = None
        _LIB = lib
    return _LIB


# (components, luma h, luma v); chroma is 1x1
SAMPLING = {"420": (3, 2, 2), "422": (3, 2, 1), "444": (3, 1, 1), "gray": (1, 1, 1), "411": (3, 4, 1),
            "440": (3, 1, 2), "410": (3, 4, 2)}


@dataclass
class ImageSpec:
    """One synthetic image: geometry, sampling, quality and content knobs."""
    width: int
    height: int
    sampling: str = "420"          # a key of SAMPLING
    quality: int = 75
    sigma: float = 6.0
    seed: int = 0
    random_tile_frac: float = 0.0
    amp_total: float = 80.0
    restart_interval: int = 0
    qtables: int = 2               # YCbCr: 2 (Y | CbCr) or 3 (Y | Cb | Cr) quant tables
    huff: int = 0                  # AC Huffman tables: 0 Annex K, 1 long codes (60 of 11 bits)
    dqt16: int = 0                 # 1: 16-bit quant tables (Pq = 1), entries up to 65535
    noninterleaved: int = 0        # 1 (YCbCr): one scan per component (table 0 redefined between scans)

    def _params(self) -> _Params:
        ncomp, h, v = SAMPLING[self.sampling]
        return _Params(self.seed, self.width, self.height, ncomp, h, v, self.quality,
                       self.restart_interval, self.sigma, self.amp_total, self.random_tile_frac,
                       self.qtables, self.huff, self.dqt16, self.noninterleaved)

    @property
    def ncomp(self) -> int:
        return SAMPLING[self.sampling][0]

    @property
    def pixels(self) -> int:
        return self.width * self.height


@dataclass
class Encoded:
    spec: ImageSpec
    data: bytes
    coefs: np.ndarray | None = None        # int16 [nblocks, 64], natural order, MCU order
    block_bits: np.ndarray | None = None   # uint32 [nblocks] (interleaved scans)
    rgb: np.ndarray | None = None          # source pixels (before encoding)


def num_blocks(spec: ImageSpec) -> int:
    return int(_lib().synth_num_blocks(ctypes.byref(spec._params())))


def quant_table(quality: int, chroma: bool) -> np.ndarray:
    out = np.zeros(64, np.uint8)
    _lib().synth_quant_table(quality, int(chroma), out.ctypes.data)
    return out.reshape(8, 8)


def source_pixels(spec: ImageSpec) -> np.ndarray:
    p = spec._params()
    out = np.empty((spec.height, spec.width, spec.ncomp), np.uint8)
    rc = _lib().synth_pixels(ctypes.byref(p), out.ctypes.data)
    if rc:
        raise RuntimeError(f"synth_pixels failed ({rc})")
    return out


def encode(spec: ImageSpec, truth: bool = False, keep_pixels: bool = False) -> Encoded:
    """Generate and encode one image; with ``truth`` also return the encoder's
    quantised coefficients and per-block bit offsets."""
    lib = _lib()
    p = spec._params()
    pix = source_pixels(spec)
    cap = spec.width * spec.height * spec.ncomp * 3 + 4096
    buf = np.empty(cap, np.uint8)
    coefs = bb = None
    if truth:
        nb = num_blocks(spec)
        coefs = np.empty((nb, 64), np.int16)
        bb = np.empty(nb, np.uint32)
    n = lib.synth_encode(ctypes.byref(p), pix.ctypes.data, buf.ctypes.data, cap,
                         coefs.ctypes.data if truth else None, bb.ctypes.data if truth else None)
    if n < 0:
        raise RuntimeError(f"synth_encode failed ({n})")
    return Encoded(spec, buf[:n].tobytes(), coefs, bb, pix if keep_pixels else None)


def encode_many(specs, truth: bool = False, threads: int | None = None):
    """Encode many images in parallel (the C calls release the GIL)."""
    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda s: encode(s, truth=truth), specs))


# --------------------------------------------------------------- workloads --
def config_plain_512(seed: int = 1000) -> ImageSpec:
    """512x512 4:2:0 q75, noise 6."""
    return ImageSpec(512, 512, "420", 75, 6.0, seed)


def config_4k(seed: int = 2000) -> ImageSpec:
    """3840x2160 4:2:0 q90, noise 12, 10% random 64x64 tiles."""
    return ImageSpec(3840, 2160, "420", 90, 12.0, seed, random_tile_frac=0.10)


def config_shapes(seed: int = 3000):
    """1024x768 gray q50, 2048x1536 4:4:4 q95, 1918x1078 4:2:0 q85; noise 8."""
    return [ImageSpec(1024, 768, "gray", 50, 8.0, seed),
            ImageSpec(2048, 1536, "444", 95, 8.0, seed + 1),
            ImageSpec(1918, 1078, "420", 85, 8.0, seed + 2)]


_ASPECTS = [(1, 1), (4, 3), (16, 9)]


def mixed_spec(rng: np.random.Generator, seed: int) -> ImageSpec:
    """One draw of the mixed distribution: area log-uniform in
    [256*256, 4096*3072], aspect 1:1/4:3/16:9, sampling 70% 4:2:0 / 20% 4:4:4 /
    10% gray, quality U{60..95}, noise U[2,12]."""
    area = math.exp(rng.uniform(math.log(256 * 256), math.log(4096 * 3072)))
    aw, ah = _ASPECTS[int(rng.integers(0, 3))]
    h = max(16, int(round(math.sqrt(area * ah / aw))))
    w = max(16, int(round(h * aw / ah)))
    u = rng.uniform()
    sampling = "420" if u < 0.7 else ("444" if u < 0.9 else "gray")
    q = int(rng.integers(60, 96))
    sigma = float(rng.uniform(2.0, 12.0))
    return ImageSpec(w, h, sampling, q, sigma, seed)


def config_mixed(n: int = 256, seed: int = 4000):
    rng = np.random.default_rng(seed)
    return [mixed_spec(rng, seed * 1000 + i) for i in range(n)]
