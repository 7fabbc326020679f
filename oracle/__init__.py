"""CPU oracle for the baseline-JPEG scan-decode path -- TEST INFRASTRUCTURE.

ctypes wrappers over ``liboracle.so`` (the project's C restatement of the
reference algorithm, ``jpeg_oracle.c``) and, where they were built, over the
reference sources compiled into ``_ref/`` (``libcpuonly.so`` = the reference
CPUonly decoder; ``libscanref.so`` = the reference ``src/decodeScanCPU.cu``).

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this package, and only as the checker.  The product decoder in
``cudajpeg_amd`` never calls it.

Parity status of each function (how it is pinned):
  parse / decode_scan   encoder ground truth (coefficients, block bit offsets)
                        + reference decodeScanCPU.cu on the golden files;
                        progressive scans: libjpeg (Pillow) progressive and
                        baseline files of the same pixels decode to the same
                        coefficients (tests/golden/progressive.npz)
  idct_int / colour     reference CPUonly pixels on the golden files
  idct_int_noshortcut   shortcut algebra (equal to idct_int on random blocks)
  idct_f64              T.81 A.3.3 definition; checked against numpy's matrix
                        form and within 1 of the integer iDCT on real blocks
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

ERROR_NAMES = {0: "NO_ERROR", 1: "SYNTAX_ERROR", 2: "UNSUPPORTED_ERROR", 3: "OOM_ERROR",
               4: "CUDA_MEM_ERROR", 5: "FILE_ERROR", 6: "CUDA_KERNEL_LAUNCH_ERROR",
               -1: "PROGRAMMER_ERROR"}


class _Channel(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("id", "samples_x", "samples_y", "dq_id", "dc_id", "ac_id",
                 "width", "height", "stride", "plane_height")]


class _Scan(ctypes.Structure):
    _fields_ = [("comp", ctypes.c_int), ("dc_id", ctypes.c_int), ("ac_id", ctypes.c_int),
                ("restart_interval", ctypes.c_int), ("scan", ctypes.c_void_p), ("scan_len", ctypes.c_size_t),
                ("bits", (ctypes.c_uint8 * 16) * 4), ("vals", (ctypes.c_uint8 * 256) * 4),
                ("ncomp", ctypes.c_int), ("comps", ctypes.c_int * 3), ("dcs", ctypes.c_int * 3),
                ("acs", ctypes.c_int * 3), ("ss", ctypes.c_int), ("se", ctypes.c_int), ("ah", ctypes.c_int),
                ("al", ctypes.c_int)]


MAX_SCANS = 64  # ORC_MAX_SCANS


class _Jpeg(ctypes.Structure):
    _fields_ = [
        ("error", ctypes.c_int), ("error_line", ctypes.c_int),
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("num_channels", ctypes.c_int),
        ("num_blocks_x", ctypes.c_int), ("num_blocks_y", ctypes.c_int),
        ("block_size_x", ctypes.c_int), ("block_size_y", ctypes.c_int),
        ("restart_interval", ctypes.c_int), ("blocks_per_mcu", ctypes.c_int),
        ("num_blocks", ctypes.c_long),
        ("ch", _Channel * 3),
        ("dq", (ctypes.c_uint8 * 64) * 4),
        ("dht_bits", (ctypes.c_uint8 * 16) * 4),
        ("dht_vals", (ctypes.c_uint8 * 256) * 4),
        ("dht_present", ctypes.c_int * 4),
        ("scan", ctypes.c_void_p), ("scan_len", ctypes.c_size_t), ("have_scan", ctypes.c_int),
        ("ext", ctypes.c_int), ("dq16", (ctypes.c_uint16 * 64) * 4),
        ("num_scans", ctypes.c_int), ("scans", _Scan * MAX_SCANS), ("comp_dq16", (ctypes.c_uint16 * 64) * 3),
        ("progressive", ctypes.c_int),
    ]


_lib = None


def build(ref: bool = True) -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref/."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference/CPUonly"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_Jpeg)
        L.orc_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t, P]
        L.orc_parse_ext.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, P]
        L.orc_decode_scan.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_planes.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_colour.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_unstuff.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_unstuff.restype = ctypes.c_size_t
        for f in ("orc_idct_int", "orc_idct_int_noshortcut", "orc_idct_f64"):
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_idct_f64_raw.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_decode_file.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_decode_file_ext.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
        _lib = L
    return _lib


@dataclass
class Parsed:
    error: int
    width: int = 0
    height: int = 0
    num_channels: int = 0
    mcus_x: int = 0
    mcus_y: int = 0
    restart_interval: int = 0
    num_blocks: int = 0
    channels: list | None = None   # dicts: samples_x/y, stride, plane_height, width, height, dq_id
    dq: np.ndarray | None = None   # [4, 64] zigzag order
    _raw: _Jpeg | None = None
    _data: bytes | None = None


EXT_DQT16 = 1  # ORC_EXT_DQT16: 16-bit quantisation tables (beyond the reference)
EXT_NONINTERLEAVED = 2  # ORC_EXT_NONINTERLEAVED: one scan per component (beyond the reference)
EXT_PROGRESSIVE = 4  # ORC_EXT_PROGRESSIVE: progressive frames, T.81 G.1 (beyond the reference)


def parse(data: bytes, ext: int = 0) -> Parsed:
    j = _Jpeg()
    e = lib().orc_parse_ext(data, len(data), ext, ctypes.byref(j))
    if e:
        return Parsed(e)
    chans = [dict(samples_x=c.samples_x, samples_y=c.samples_y, stride=c.stride,
                  plane_height=c.plane_height, width=c.width, height=c.height, dq_id=c.dq_id,
                  dc_id=c.dc_id, ac_id=c.ac_id)
             for c in list(j.ch)[: j.num_channels]]
    dq = np.array([list(t) for t in j.dq16], np.uint16)
    return Parsed(0, j.width, j.height, j.num_channels, j.num_blocks_x, j.num_blocks_y,
                  j.restart_interval, j.num_blocks, chans, dq, j, data)


@dataclass
class ScanResult:
    error: int
    qcoef: np.ndarray | None = None      # int16 [nblocks, 64] natural order
    bitpos: np.ndarray | None = None     # uint32 [nblocks]
    work: list | None = None             # per component int32 planes (reference layout)


def decode_scan(p: Parsed, want_work: bool = False) -> ScanResult:
    q = np.zeros((p.num_blocks, 64), np.int16)
    bp = np.zeros(p.num_blocks, np.uint32)
    work = None
    wptr = None
    if want_work:
        work = [np.zeros(c["stride"] * c["plane_height"], np.int32) for c in p.channels]
        arr = (ctypes.c_void_p * 3)(*([w.ctypes.data for w in work] + [None] * (3 - len(work))))
        wptr = ctypes.cast(arr, ctypes.c_void_p)
    e = lib().orc_decode_scan(ctypes.byref(p._raw), q.ctypes.data, wptr, bp.ctypes.data)
    return ScanResult(e, q, bp, work)


def planes(p: Parsed, qcoef: np.ndarray, mode: int = 0) -> list:
    """Planar component planes (stride x plane_height); mode 0 = integer iDCT,
    1 = FP64 iDCT."""
    pl = [np.zeros((c["plane_height"], c["stride"]), np.uint8) for c in p.channels]
    arr = (ctypes.c_void_p * 3)(*([x.ctypes.data for x in pl] + [None] * (3 - len(pl))))
    qc = np.ascontiguousarray(qcoef, np.int16)
    lib().orc_planes(ctypes.byref(p._raw), qc.ctypes.data, ctypes.cast(arr, ctypes.c_void_p), mode)
    return pl


def colour(p: Parsed, pl: list) -> np.ndarray:
    out = np.zeros((p.height, p.width, 3 if p.num_channels == 3 else 1), np.uint8)
    pls = [np.ascontiguousarray(x) for x in pl]
    arr = (ctypes.c_void_p * 3)(*([x.ctypes.data for x in pls] + [None] * (3 - len(pls))))
    lib().orc_colour(ctypes.byref(p._raw), ctypes.cast(arr, ctypes.c_void_p), out.ctypes.data)
    return out


def decode(data: bytes, mode: int = 0, ext: int = 0):
    """Whole decode: returns (error, HxWx{1,3} uint8)."""
    w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = parse(data, ext)
    if p.error:
        return p.error, None
    out = np.zeros(p.width * p.height * p.num_channels, np.uint8)
    e = lib().orc_decode_file_ext(data, len(data), ext, mode, out.ctypes.data, out.size,
                                  ctypes.byref(w), ctypes.byref(h), ctypes.byref(c))
    if e:
        return e, None
    return 0, out.reshape(h.value, w.value, c.value)


def unstuff(scan: bytes) -> bytes:
    out = ctypes.create_string_buffer(len(scan) + 2)
    n = lib().orc_unstuff(scan, len(scan), out)
    return out.raw[:n]


def scan_bytes(p: Parsed) -> bytes:
    """The stuffed entropy-coded segment of a parsed file."""
    base = ctypes.cast(ctypes.c_char_p(p._data), ctypes.c_void_p).value
    off = p._raw.scan - base
    return p._data[off: off + p._raw.scan_len]


def idct(block_deq: np.ndarray, kind: str = "int") -> np.ndarray:
    b = np.ascontiguousarray(block_deq, np.int32).reshape(64)
    out = np.zeros(64, np.uint8)
    f = {"int": "orc_idct_int", "int_noshortcut": "orc_idct_int_noshortcut", "f64": "orc_idct_f64"}[kind]
    getattr(lib(), f)(b.ctypes.data, out.ctypes.data, 8)
    return out.reshape(8, 8)


def idct_f64_raw(block_deq: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(block_deq, np.int32).reshape(64)
    out = np.zeros(64, np.float64)
    lib().orc_idct_f64_raw(b.ctypes.data, out.ctypes.data)
    return out.reshape(8, 8)


# ------------------------------------------------------ compiled reference --
_cpuonly = _scanref = None


def have_ref() -> bool:
    return (os.path.exists(os.path.join(REF_DIR, "libcpuonly.so"))
            and os.path.exists(os.path.join(REF_DIR, "libscanref.so")))


def ref_cpuonly():
    global _cpuonly
    if _cpuonly is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libcpuonly.so"))
        L.ref_cpuonly_decode.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_long,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
        L.ref_cpuonly_time.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]
        _cpuonly = L
    return _cpuonly


def ref_cpuonly_decode(path: str, cap: int):
    out = np.zeros(cap, np.uint8)
    w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    e = ref_cpuonly().ref_cpuonly_decode(path.encode(), out.ctypes.data, cap, ctypes.byref(w),
                                         ctypes.byref(h), ctypes.byref(c))
    if e:
        return e, None
    return 0, out[: w.value * h.value * c.value].reshape(h.value, w.value, c.value)


def ref_cpuonly_time(paths, reps: int = 1) -> np.ndarray:
    arr = (ctypes.c_char_p * len(paths))(*[p.encode() for p in paths])
    t = np.zeros(len(paths) * reps, np.float64)
    fails = ref_cpuonly().ref_cpuonly_time(arr, len(paths), reps, t.ctypes.data)
    if fails:
        raise RuntimeError(f"reference CPUonly failed on {fails} decodes")
    return t.reshape(len(paths), reps)


def ref_scan_decode(data: bytes, p: Parsed):
    global _scanref
    if _scanref is None:
        _scanref = ctypes.CDLL(os.path.join(REF_DIR, "libscanref.so"))
        _scanref.ref_scan_decode.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_void_p]
    work = [np.zeros(c["stride"] * c["plane_height"], np.int32) for c in p.channels]
    arr = (ctypes.c_void_p * 3)(*([w.ctypes.data for w in work] + [None] * (3 - len(work))))
    e = _scanref.ref_scan_decode(data, len(data), ctypes.cast(arr, ctypes.c_void_p))
    return e, work
