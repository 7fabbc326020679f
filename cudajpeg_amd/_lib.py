"""ctypes binding of the C ABI in include/cudajpeg_amd.h.

Loads the in-tree libcudajpeg_amd.so (built by __graft_entry__.build()).  There
is no fallback: if the library is missing, importing the decoder fails.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcudajpeg_amd.so")

NO_ERROR, SYNTAX_ERROR, UNSUPPORTED_ERROR, OOM_ERROR = 0, 1, 2, 3
DEVICE_MEM_ERROR, FILE_ERROR, KERNEL_LAUNCH_ERROR, PROGRAMMER_ERROR = 4, 5, 6, -1
IDCT_INT, IDCT_FLOAT, IDCT_MFMA = 0, 1, 2
ENTROPY_TILED, ENTROPY_SERIAL = 0, 1
UNSTUFF_HOST, UNSTUFF_DEVICE = 0, 1
EXT_DQT16 = 1
EXT_NONINTERLEAVED = 2
EXT_PROGRESSIVE = 4
OUT_PIXELS, OUT_COEFFICIENTS, OUT_PLANE0 = 0, 1, 2

# Every symbol include/cudajpeg_amd.h and include/format.h declare.
EXPORTS = [
    "cjd_parse", "cjd_parse_ext", "cjd_context_create", "cjd_context_destroy", "cjd_batch_create", "cjd_batch_destroy",
    "cjd_batch_set_options", "cjd_batch_set_tuning", "cjd_batch_set_overlap", "cjd_batch_set_unstuff", "cjd_batch_set_extensions", "cjd_batch_set_tune_table", "cjd_batch_stage", "cjd_batch_stage_async", "cjd_batch_decode", "cjd_batch_decode_scan",
    "cjd_batch_transform", "cjd_batch_fallbacks", "cjd_batch_set_wait_limit", "cjd_batch_protocol_redos",
    "cjd_batch_status",
    "cjd_batch_finish", "cjd_batch_capture",
    "cjd_batch_replay", "cjd_batch_timings", "cjd_batch_kernel_timings", "cjd_batch_image_info", "cjd_batch_device_ptr",
    "cjd_batch_download", "cjd_batch_totals", "cjd_decode_noalloc", "cjd_noalloc_scratch_bytes", "cjd_error_name",
    "cjd_server_create", "cjd_server_submit", "cjd_server_poll", "cjd_server_release", "cjd_server_stats", "cjd_server_now_ns",
    "cjd_server_replay", "cjd_server_destroy",
    "newJPGReader", "delJPGReader", "openJPG", "openJPGMem", "writeJPG", "printError", "read16", "skipBlock",
    "decodeSOF", "decodeDHT", "decodeDQT", "decodeDRI", "decodeScanGPU", "decodeScanCPU",
    "iDCT_resample_colourTransform",
]


class FrameInfo(ctypes.Structure):
    """cjd_frame_info"""
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("num_channels", ctypes.c_int32),
        ("mcus_x", ctypes.c_int32), ("mcus_y", ctypes.c_int32),
        ("mcu_width", ctypes.c_int32), ("mcu_height", ctypes.c_int32),
        ("restart_interval", ctypes.c_int32), ("blocks_per_mcu", ctypes.c_int32),
        ("num_blocks", ctypes.c_int64),
        ("comp_id", ctypes.c_int32 * 3), ("samples_x", ctypes.c_int32 * 3), ("samples_y", ctypes.c_int32 * 3),
        ("comp_width", ctypes.c_int32 * 3), ("comp_height", ctypes.c_int32 * 3),
        ("comp_stride", ctypes.c_int32 * 3), ("comp_plane_height", ctypes.c_int32 * 3),
        ("dq_id", ctypes.c_int32 * 3), ("dc_id", ctypes.c_int32 * 3), ("ac_id", ctypes.c_int32 * 3),
        ("dq_tables", (ctypes.c_uint8 * 64) * 4),
        ("huff_bits", (ctypes.c_uint8 * 16) * 4),
        ("huff_vals", (ctypes.c_uint8 * 256) * 4),
        ("scan_offset", ctypes.c_uint64), ("scan_length", ctypes.c_uint64),
        ("error_line", ctypes.c_int32), ("num_scans", ctypes.c_int32),
        ("dq_tables16", (ctypes.c_uint16 * 64) * 4),
        ("progressive", ctypes.c_int32), ("reserved_", ctypes.c_int32),
    ]


class TuneEntry(ctypes.Structure):
    """cjd_tune_entry"""
    _fields_ = [("max_scan_bytes", ctypes.c_uint64), ("tile_bits", ctypes.c_int), ("ck_bits", ctypes.c_int),
                ("item_tiles", ctypes.c_int), ("max_bits_per_block", ctypes.c_int)]


class ServerOptions(ctypes.Structure):
    """cjd_server_options"""
    _fields_ = [("max_batch", ctypes.c_int), ("max_wait_us", ctypes.c_int), ("frame_budget", ctypes.c_uint64),
                ("slots", ctypes.c_int), ("idct_mode", ctypes.c_int), ("tile_bits", ctypes.c_int),
                ("extensions", ctypes.c_int)]


class ServerCounters(ctypes.Structure):
    """cjd_server_counters"""
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "pending", "inflight", "redone", "batches",
                                              "stage_ns", "launch_ns", "finish_ns", "slot_wait_ns", "device_ns",
                                              "device_max_ns", "turnaround_ns", "finish_max_ns",
                                              "protocol_redos")]


class Completion(ctypes.Structure):
    """cjd_completion"""
    _fields_ = [("id", ctypes.c_uint64), ("error", ctypes.c_int), ("batch_size", ctypes.c_int),
                ("frame", ctypes.c_void_p), ("frame_bytes", ctypes.c_uint64),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("channels", ctypes.c_int),
                ("t_arrival_ns", ctypes.c_int64), ("t_launch_ns", ctypes.c_int64), ("t_ready_ns", ctypes.c_int64)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    FP = ctypes.POINTER(FrameInfo)
    sig = {
        "cjd_parse": ([ctypes.c_char_p, sz, FP], i32),
        "cjd_parse_ext": ([ctypes.c_char_p, sz, i32, FP], i32),
        "cjd_context_create": ([i32, ctypes.POINTER(vp)], i32),
        "cjd_context_destroy": ([vp], None),
        "cjd_batch_create": ([vp, ctypes.POINTER(vp)], i32),
        "cjd_batch_destroy": ([vp], None),
        "cjd_batch_set_options": ([vp, i32, i32, i32, i32], i32),
        "cjd_batch_set_tuning": ([vp, i32, i32], i32),
        "cjd_batch_set_overlap": ([vp, i32], i32),
        "cjd_batch_set_unstuff": ([vp, i32], i32),
        "cjd_batch_set_extensions": ([vp, i32], i32),
        "cjd_batch_set_tune_table": ([vp, i32, ctypes.POINTER(TuneEntry)], i32),
        "cjd_batch_stage": ([vp, i32, vp, vp, vp], i32),
        "cjd_batch_stage_async": ([vp, i32, vp, vp, vp, vp], i32),
        "cjd_batch_decode": ([vp, vp], i32),
        "cjd_batch_decode_scan": ([vp, vp], i32),
        "cjd_batch_transform": ([vp, vp], i32),
        "cjd_batch_fallbacks": ([vp], i32),
        "cjd_batch_set_wait_limit": ([vp, ctypes.c_int64], i32),
        "cjd_batch_protocol_redos": ([vp], i32),
        "cjd_batch_status": ([vp, i32, ctypes.POINTER(ctypes.c_uint32)], i32),
        "cjd_batch_finish": ([vp, vp, vp], i32),
        "cjd_batch_capture": ([vp, vp], i32),
        "cjd_batch_replay": ([vp, vp], i32),
        "cjd_batch_timings": ([vp, vp, vp, vp], i32),
        "cjd_batch_kernel_timings": ([vp, vp, vp], i32),
        "cjd_batch_image_info": ([vp, i32, FP], i32),
        "cjd_batch_device_ptr": ([vp, i32, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)], i32),
        "cjd_batch_download": ([vp, i32, i32, vp, sz], i32),
        "cjd_batch_totals": ([vp, vp, vp, vp, vp, vp], i32),
        "cjd_decode_noalloc": ([i32, ctypes.c_char_p, sz, i32, vp, sz, vp, ctypes.POINTER(sz)], i32),
        "cjd_noalloc_scratch_bytes": ([], sz),
        "cjd_error_name": ([i32], ctypes.c_char_p),
        "cjd_server_create": ([vp, ctypes.POINTER(ServerOptions), ctypes.POINTER(vp)], i32),
        "cjd_server_submit": ([vp, ctypes.c_char_p, sz, ctypes.POINTER(ctypes.c_uint64)], i32),
        "cjd_server_poll": ([vp, ctypes.POINTER(Completion), i32, i32], i32),
        "cjd_server_release": ([vp, ctypes.c_uint64], i32),
        "cjd_server_stats": ([vp, ctypes.POINTER(ServerCounters)], i32),
        "cjd_server_now_ns": ([], ctypes.c_int64),
        "cjd_server_replay": ([vp, i32, vp, vp, vp, vp, vp, vp], i32),
        "cjd_server_destroy": ([vp], None),
        "newJPGReader": ([], vp),
        "delJPGReader": ([vp], None),
        "openJPG": ([vp, ctypes.c_char_p], i32),
        "openJPGMem": ([vp, ctypes.c_char_p, sz], i32),
        "writeJPG": ([vp, ctypes.c_char_p], None),
        "printError": ([vp], None),
        "decodeScanGPU": ([vp], None),
        "decodeScanCPU": ([vp], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def error_name(code: int) -> str:
    return lib().cjd_error_name(code).decode()
