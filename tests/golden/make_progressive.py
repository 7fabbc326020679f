"""Writes progressive.npz: pairs of JPEG files of the same pixels, one baseline
(SOF0) and one progressive (SOF2), both written by libjpeg through Pillow with
the same quality and chroma subsampling.

libjpeg quantises both from the same forward DCT, so the two files hold the
same quantised coefficients (the blocks by which MCUs pad a component's grid
included: both take the last real block's DC and zero AC there).  The
baseline file's decode is pinned by the reference (tests/golden/golden.npz,
tests/test_oracle.py); the pair pins the progressive decode.  Pillow's
progressive files use libjpeg's standard script (jcparam.c
jpeg_simple_progression): DC first (Al 1) over all components, AC first
bands of luma (1-5 and 6-63, Al 2) and chroma (1-63, Al 1), an AC refinement
of luma (Ah 2, Al 1), the DC refinement and AC refinements down to Al 0, with
Huffman tables optimised per scan (a DHT before every scan), and, with
restart markers, a restart interval per scan.

Run in the build container:  python tests/golden/make_progressive.py
"""
import io
import os

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))


def pixels(w, h, seed, gray=False):
    """Smooth gradients, a ringing pattern, edges and noise: coefficients of
    every size, long zero runs (end-of-band runs) and bands of refinement."""
    rs = np.random.RandomState(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    r = (x * 2 + 40 * np.sin(y / 7.0)) % 256
    g = (y * 3 + x + 50 * (((x // 16) + (y // 16)) % 2)) % 256
    b = 128 + 60 * np.cos((x + y) / 9.0)
    img = np.stack([r, g, b], -1) + rs.randn(h, w, 3) * rs.choice([0.5, 4.0, 16.0], (h, w, 1))
    flat = rs.rand(h // 16 + 1, w // 16 + 1) < 0.3  # flat 16x16 areas: all-zero AC blocks
    img[np.kron(flat, np.ones((16, 16), bool))[:h, :w]] = 90.0
    img = img.clip(0, 255).astype(np.uint8)
    return img[..., 0] if gray else img


# name: (width, height, seed, gray, quality, save options)
CASES = {
    "c420_64x48_q75": (64, 48, 1, False, 75, dict(subsampling=2)),
    "c420_33x17_q90": (33, 17, 2, False, 90, dict(subsampling=2)),
    "c444_40x24_q95": (40, 24, 3, False, 95, dict(subsampling=0)),
    "c422_37x29_q80": (37, 29, 4, False, 80, dict(subsampling=1)),
    "gray_48x32_q50": (48, 32, 5, True, 50, {}),
    "gray_19x45_q100": (19, 45, 6, True, 100, {}),
    "c420_131x77_q85_rstb3": (131, 77, 7, False, 85, dict(subsampling=2, restart_marker_blocks=3)),
    "c422_96x40_q92_rstr1": (96, 40, 8, False, 92, dict(subsampling=1, restart_marker_rows=1)),
    "c444_50x50_q100": (50, 50, 9, False, 100, dict(subsampling=0)),
    "gray_70x9_q70_rstb1": (70, 9, 10, True, 70, dict(restart_marker_blocks=1)),
    "c420_8x8_q60": (8, 8, 11, False, 60, dict(subsampling=2)),
}


def encode(w, h, seed, gray, q, opts, progressive):
    b = io.BytesIO()
    Image.fromarray(pixels(w, h, seed, gray)).save(b, "JPEG", quality=q, progressive=progressive, **opts)
    return b.getvalue()


def main():
    out = {}
    for name, (w, h, seed, gray, q, opts) in CASES.items():
        out[name + "/base"] = np.frombuffer(encode(w, h, seed, gray, q, opts, False), np.uint8)
        out[name + "/prog"] = np.frombuffer(encode(w, h, seed, gray, q, opts, True), np.uint8)
    np.savez_compressed(os.path.join(HERE, "progressive.npz"), **out)
    print(len(CASES), "pairs,", sum(v.size for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
