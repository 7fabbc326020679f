"""Serving loop: JPEG byte streams in, device-resident frames out.

Wraps the cjd_server_* C ABI: requests are submitted as they arrive, batched
(flush at max_batch waiting requests or after max_wait_us), admitted by
decoded-frame memory, decoded on the GPU while the next batch is staged, and
handed back as completions that point at the frame in device memory.  Every
completion must be released.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .decoder import Context, _check, default_context


@dataclass
class Completion:
    id: int
    error: int
    batch_size: int
    frame: int            # device pointer (0 on error)
    frame_bytes: int
    width: int
    height: int
    channels: int
    t_arrival_ns: int
    t_launch_ns: int
    t_ready_ns: int

    @property
    def latency_ms(self) -> float:
        return (self.t_ready_ns - self.t_arrival_ns) / 1e6


class Server:
    def __init__(self, max_batch: int = 64, max_wait_us: int = 2000, frame_budget: int = 64 << 30, slots: int = 4,
                 idct: int = L.IDCT_MFMA, tile_bits: int = 0, ctx: Context | None = None, extensions: int = 0):
        self.ctx = ctx or default_context()
        opt = L.ServerOptions(max_batch, max_wait_us, frame_budget, slots, idct, tile_bits, extensions)
        self.h = ctypes.c_void_p()
        _check(L.lib().cjd_server_create(self.ctx.h, ctypes.byref(opt), ctypes.byref(self.h)), "server: ")
        self._buf = (L.Completion * 256)()

    def submit(self, data: bytes) -> int:
        rid = ctypes.c_uint64()
        _check(L.lib().cjd_server_submit(self.h, data, len(data), ctypes.byref(rid)), "submit: ")
        return rid.value

    def poll(self, max_n: int = 256, wait_us: int = 0) -> list[Completion]:
        n = L.lib().cjd_server_poll(self.h, self._buf, min(max_n, len(self._buf)), wait_us)
        out = []
        for i in range(max(n, 0)):
            c = self._buf[i]
            out.append(Completion(c.id, c.error, c.batch_size, c.frame or 0, c.frame_bytes, c.width, c.height,
                                  c.channels, c.t_arrival_ns, c.t_launch_ns, c.t_ready_ns))
        return out

    def release(self, rid: int) -> None:
        _check(L.lib().cjd_server_release(self.h, rid), "release: ")

    def stats(self) -> dict:
        """Counters of the loop (cjd_server_counters): requests submitted,
        completed, pending and in flight, images redone, batches, and host
        ns spent staging, launching, finishing and waiting for a slot."""
        c = L.ServerCounters()
        _check(L.lib().cjd_server_stats(self.h, ctypes.byref(c)), "stats: ")
        return {k: getattr(c, k) for k, _ in L.ServerCounters._fields_}

    def frame_tensor(self, c: Completion):
        """Zero-copy torch view (uint8, H x W x channels) of a completed
        frame in device memory; valid until the request is released."""
        return device_frame(c.frame, c.height, c.width, c.channels)

    def download(self, c: Completion) -> np.ndarray:
        """Copy a completed frame to host memory."""
        return self.frame_tensor(c).cpu().numpy()

    def replay(self, files: list[bytes], arrival_ns: np.ndarray, progress=None, every_s: float = 2.0):
        """Submit file k at start + arrival_ns[k] from a native thread, release
        each frame as it completes.  Returns (latency_ns, errors, wall_ns).
        progress(stats) is called every every_s seconds while it runs."""
        n = len(files)
        keep = [ctypes.create_string_buffer(f, len(f)) for f in files]
        ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(b, ctypes.c_void_p) for b in keep])
        sizes = (ctypes.c_size_t * n)(*[len(f) for f in files])
        arr = np.ascontiguousarray(arrival_ns, np.int64)
        lat = np.zeros(n, np.int64)
        err = np.zeros(n, np.int32)
        wall = ctypes.c_int64()
        rc = [0]

        def run():
            rc[0] = L.lib().cjd_server_replay(self.h, n, ptrs, sizes, arr.ctypes.data, lat.ctypes.data,
                                              err.ctypes.data, ctypes.byref(wall))

        if progress is None:
            run()
        else:
            import threading
            t = threading.Thread(target=run)
            t.start()
            while True:
                t.join(every_s)
                if not t.is_alive():
                    break
                progress(self.stats())
        _check(rc[0], "replay: ")
        return lat, err, wall.value

    def close(self) -> None:
        if self.h:
            L.lib().cjd_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceArray:
    """__cuda_array_interface__ (v3) of a uint8 frame in device memory; torch
    builds a tensor on it without copying."""

    def __init__(self, ptr: int, shape: tuple):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": "|u1", "data": (int(ptr), False),
                                         "version": 3, "strides": None, "stream": None}


def device_frame(ptr: int, height: int, width: int, channels: int):
    """torch.uint8 view of a decoded frame at device address ptr."""
    import torch
    return torch.as_tensor(_DeviceArray(ptr, (height, width, channels)), device="cuda")
