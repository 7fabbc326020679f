"""Writes the golden fixtures from the REFERENCE's own code (run in the build
container, where /root/reference exists and oracle/_ref has been built from it).

For each tiny synthetic JPEG (the project's seeded encoder) it stores:
  * the file bytes,
  * the dequantised working_space per component produced by the reference
    src/decodeScanCPU.cu (compiled into oracle/_ref/libscanref.so),
  * the RGB / gray pixels produced by the reference CPUonly decoder
    (oracle/_ref/libcpuonly.so), when CPUonly decodes the file.
Only data is committed: golden.npz (+ this script).
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cudajpeg_amd import synth  # noqa: E402

CASES = [
    ("c420_64x48_q75", synth.ImageSpec(64, 48, "420", 75, 6.0, 101)),
    ("c420_33x17_q90", synth.ImageSpec(33, 17, "420", 90, 8.0, 102)),
    ("c444_40x24_q95", synth.ImageSpec(40, 24, "444", 95, 8.0, 103)),
    ("gray_48x32_q50", synth.ImageSpec(48, 32, "gray", 50, 8.0, 104)),
    ("c422_37x29_q80", synth.ImageSpec(37, 29, "422", 80, 5.0, 105)),
    ("c420_96x64_q85_rst2", synth.ImageSpec(96, 64, "420", 85, 6.0, 106, restart_interval=2)),
    ("c420_80x80_q95_noise", synth.ImageSpec(80, 80, "420", 95, 12.0, 107, random_tile_frac=0.5)),
    ("gray_8x8_q100", synth.ImageSpec(8, 8, "gray", 100, 2.0, 108)),
    ("c411_72x40_q85", synth.ImageSpec(72, 40, "411", 85, 6.0, 109)),
    ("c440_41x35_q70", synth.ImageSpec(41, 35, "440", 70, 6.0, 110)),
    ("c410_70x34_q90", synth.ImageSpec(70, 34, "410", 90, 8.0, 111)),
]


def main():
    if not oracle.have_ref():
        raise SystemExit("oracle/_ref is not built (needs /root/reference): make -C oracle ref")
    out = {}
    for name, spec in CASES:
        enc = synth.encode(spec)
        p = oracle.parse(enc.data)
        e, work = oracle.ref_scan_decode(enc.data, p)
        assert e == 0, (name, e)
        out[f"{name}/file"] = np.frombuffer(enc.data, np.uint8)
        for c, w in enumerate(work):
            out[f"{name}/work{c}"] = w
        with tempfile.NamedTemporaryFile(suffix=".jpg", dir=HERE, delete=False) as f:
            f.write(enc.data)
            path = f.name
        e, pix = oracle.ref_cpuonly_decode(path, spec.width * spec.height * 3)
        os.unlink(path)
        if e == 0:
            out[f"{name}/pixels"] = pix
        print(name, len(enc.data), "bytes", "cpuonly" if e == 0 else f"cpuonly failed ({e})")
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
