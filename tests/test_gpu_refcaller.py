"""The reader boundary on the GPU, through the reference's own callers.

* The reference's host program (src/main.cu, built by tests/refcaller/Makefile
  against include/format.h) opens files with openJPG, writes outfile.ppm /
  outfile.pgm with writeJPG and reads reader->time: its frame equals the
  oracle's pixels and it prints its two timing lines.
* decodeScanCPU (include/decodeScanCPU.h:6) called at the SOS marker, as
  CPUonly's marker loop does (CPUonly/src/main.c:34), fills working_space
  exactly as the reference's own decodeScanCPU did for the golden files
  (tests/golden, written by src/decodeScanCPU.cu)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from cudajpeg_amd import _lib as L
from cudajpeg_amd import JPGReader, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "refcaller", "_bin", "ref_main")
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def _read_pnm(path):
    data = open(path, "rb").read()
    tokens, pos = [], 0
    while len(tokens) < 4:  # magic, width, height, maxval
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        tokens.append(data[pos:end])
        pos = end
    pos += 1
    w, h = int(tokens[1]), int(tokens[2])
    return np.frombuffer(data[pos:], np.uint8).reshape(h, w, 3 if tokens[0] == b"P6" else 1)


@pytest.mark.skipif(not os.path.exists(BIN), reason="reference caller not built (needs /root/reference at build)")
@pytest.mark.parametrize("spec", [synth.ImageSpec(96, 64, "420", 80, 6.0, 7),
                                  synth.ImageSpec(45, 31, "gray", 70, 6.0, 8),
                                  synth.ImageSpec(640, 480, "422", 85, 8.0, 9, restart_interval=5)])
def test_reference_main_decodes_through_the_reader_api(gpu, tmp_path, spec):
    enc = synth.encode(spec, truth=True)
    path = tmp_path / "in.jpg"
    path.write_bytes(enc.data)
    print("case ref_main", spec, flush=True)
    out = subprocess.run([BIN, str(path)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    assert "ms per image" in out.stdout and "DEBUG_TIME" in out.stdout, out.stdout
    name = "outfile.pgm" if spec.sampling == "gray" else "outfile.ppm"
    p = oracle.parse(enc.data)
    ref = oracle.colour(p, oracle.planes(p, enc.coefs, 0))
    assert np.array_equal(_read_pnm(tmp_path / name), ref)


def _sos_offset(data: bytes) -> int:
    """Offset of the SOS marker, walking the marker segments."""
    pos = 2
    while data[pos + 1] != 0xDA:
        pos += 2 + ((data[pos + 2] << 8) | data[pos + 3])
    return pos


def test_decode_scan_cpu_equals_reference_working_space(gpu):
    g = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in g.files})
    r = JPGReader()
    lib = L.lib()
    for n in names:
        data = g[f"{n}/file"].tobytes()
        print("case decodeScanCPU", n, flush=True)
        assert r.open(data) == 0, (n, r.error)  # SOF / DHT / DQT / DRI state
        raw = r.raw
        raw.time = 0
        # at SOS, past the marker, as the marker loop calls the scan decoder
        raw.pos = ctypes.cast(raw.file_buf.mem, ctypes.c_void_p).value + _sos_offset(data) + 2
        lib.decodeScanCPU(r._p)
        assert raw.error == 0, (n, r.error)
        assert raw.pos == raw.end
        for c in range(raw.num_channels):
            assert np.array_equal(r.working_space(c), g[f"{n}/work{c}"]), (n, c)
    r.close()
