"""Autotuners of the tiled entropy decoder's parameters per image bucket.

For each bucket of sample images (by entropy-coded scan bytes) it times the
candidate (tile bits, tiles per work item) pairs on the GPU and keeps the
fastest:
  mode "latency":    each image decoded alone (one image per launch); the
                     objective is the mean decode time per image;
  mode "throughput": the bucket's images decoded as one batch; the objective
                     is time per pixel.
The result is a table for BatchDecoder.set_tune_table / cjd_batch_set_tune_table:
rows (max_scan_bytes, tile_bits, ck_bits, item_tiles), the checkpoint spacing
left at its default (tile_bits / 16).

autotune_batched tunes the batched form as a whole: buckets by scan size and
density (scan bits per 8x8 block), and the objective is the decode time of
the whole batch (all buckets share its launches), minimised by coordinate
descent over the buckets from the best uniform setting; rows
(max_scan_bytes, tile_bits, 0, item_tiles, max_bits_per_block).
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


BATCH_SIZES = (128 << 10, 512 << 10, 2 << 20, (1 << 63) - 1)
BATCH_DENSITIES = (40, 96, 0)  # bits per block; 0 = any
BATCH_CANDIDATES = tuple((s, it) for s in (4096, 6144, 8192) for it in (64, 128, 256))


def _batch_time(b, files, table, reps):
    b.set_tune_table(table)
    if b.stage(files).any():
        return float("inf")
    t = []
    for _ in range(reps + 1):
        b.decode()
        b.finish()
        t.append(b.timings()[2])
    return float(np.median(t[1:]))


def autotune_batched(files: list[bytes], idct: int = L.IDCT_MFMA, parts: int = 2, reps: int = 5,
                     candidates=BATCH_CANDIDATES, sizes=BATCH_SIZES, densities=BATCH_DENSITIES, min_gain=0.003,
                     log=None) -> list[tuple]:
    """Per-bucket (tile bits, tiles per item) of the batched decode of `files`,
    measured on the whole batch; returns the tune table (buckets without
    images are left out)."""
    info = [parse(f)[1] for f in files]
    keys = []
    for f in info:
        bpb = f.scan_length * 8 / max(1, f.num_blocks)
        size = next(s for s in sizes if f.scan_length <= s)
        dens = next(d for d in densities if d == 0 or bpb <= d)
        keys.append((size, dens))
    buckets = sorted(set(keys), key=lambda k: (k[0], k[1] or 1 << 30))
    b = BatchDecoder(idct)
    b.set_overlap(parts)

    def table_of(choice):
        return [(k[0], choice[k][0], 0, choice[k][1], k[1]) for k in buckets]

    best_t, best = float("inf"), None
    for c in candidates:  # the best uniform setting first
        choice = {k: c for k in buckets}
        t = _batch_time(b, files, table_of(choice), reps)
        if log:
            log(f"autotune batched uniform {c}: {t:.3f} ms")
        if t < best_t:
            best_t, best = t, choice
    for k in buckets:  # then one bucket at a time
        for c in candidates:
            if c == best[k]:
                continue
            choice = dict(best)
            choice[k] = c
            t = _batch_time(b, files, table_of(choice), reps)
            if t < best_t * (1 - min_gain):
                if log:
                    log(f"autotune batched bucket {k} -> {c}: {t:.3f} ms (was {best_t:.3f})")
                best_t, best = t, choice
    b.close()
    return table_of(best)


def save_table(table, path: str) -> None:
    with open(path, "w") as f:
        json.dump([list(t) for t in table], f)


def load_table(path: str) -> list[tuple]:
    """A table saved by save_table, or a {"table": [...]} document."""
    with open(path) as f:
        doc = json.load(f)
    rows = doc["table"] if isinstance(doc, dict) else doc
    return [tuple(t) for t in rows]


# the batched-form table measured on MI355X (see its description)
BATCHED_TABLE_MI355X = __file__.rsplit("/", 1)[0] + "/tune/mi355x_batched.json"
