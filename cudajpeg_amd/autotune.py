"""Autotuner of the tiled entropy decoder's parameters per scan-size bucket.

For each bucket of sample images (by entropy-coded scan bytes) it times the
candidate (tile bits, tiles per work item) pairs on the GPU and keeps the
fastest:
  mode "latency":    each image decoded alone (one image per launch); the
                     objective is the mean decode time per image;
  mode "throughput": the bucket's images decoded as one batch; the objective
                     is time per pixel.
The result is a table for BatchDecoder.set_tune_table / cjd_batch_set_tune_table:
rows (max_scan_bytes, tile_bits, ck_bits, item_tiles), the checkpoint spacing
left at its default (tile_bits / 8).
"""
from __future__ import annotations

import json

import numpy as np

from . import _lib as L
from .decoder import BatchDecoder, parse

BUCKETS = (32 << 10, 128 << 10, 512 << 10, 2 << 20, 8 << 20, 1 << 62)
CANDIDATES = tuple((s, it) for s in (1024, 2048, 4096, 8192) for it in (16, 64, 256))


def _time(files, idct, S, items, reps):
    b = BatchDecoder(idct, tile_bits=S)
    b.set_tuning(0, items)
    if b.stage(files).any():
        b.close()
        return float("inf")
    b.decode()
    b.finish()
    t = []
    for _ in range(reps):
        b.decode()
        b.finish()
        t.append(b.timings()[2])
    b.close()
    return float(np.median(t))


def autotune(files: list[bytes], mode: str = "latency", idct: int = L.IDCT_FLOAT, reps: int = 3,
             per_bucket: int = 3, candidates=CANDIDATES, log=None) -> list[tuple]:
    scan = [parse(f)[1].scan_length for f in files]
    table = []
    lo = 0
    for hi in BUCKETS:
        idx = [i for i, s in enumerate(scan) if lo < s <= hi]
        lo = hi
        if not idx:
            continue
        if mode == "latency":
            idx = sorted(idx, key=lambda i: scan[i])
            pick = [idx[int(k)] for k in np.linspace(0, len(idx) - 1, min(per_bucket, len(idx)))]
        else:
            pick = idx
        best, best_t = None, float("inf")
        for S, items in candidates:
            if mode == "latency":
                t = float(np.mean([_time([files[i]], idct, S, items, reps) for i in pick]))
            else:
                t = _time([files[i] for i in pick], idct, S, items, reps)
            if log:
                log(f"autotune {mode} bucket <= {hi} ({len(pick)} images): tile {S} items {items}: {t:.3f} ms")
            if t < best_t:
                best, best_t = (S, items), t
        table.append((int(min(hi, (1 << 63) - 1)), best[0], 0, best[1]))
    return table


def save_table(table, path: str) -> None:
    with open(path, "w") as f:
        json.dump([list(t) for t in table], f)


def load_table(path: str) -> list[tuple]:
    with open(path) as f:
        return [tuple(t) for t in json.load(f)]
