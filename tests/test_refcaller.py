"""The reference's own host program (src/main.cu), built by
tests/refcaller/Makefile against include/format.h from a copy with the CUDA
runtime calls made HIP's: it links against this library and runs.  Without
an argument it prints its usage and exits with 1 before touching the reader
(src/main.cu:12-15), so this needs no GPU; tests/test_gpu_refcaller.py runs it
on files."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "refcaller", "_bin", "ref_main")


@pytest.mark.skipif(not os.path.exists(BIN), reason="reference caller not built (needs /root/reference)")
def test_reference_main_links_against_the_reader_api():
    ldd = subprocess.run(["ldd", BIN], capture_output=True, text=True, check=True).stdout
    line = [ln for ln in ldd.splitlines() if "libcudajpeg_amd.so" in ln]
    assert line and "not found" not in line[0], ldd
    assert os.path.realpath(line[0].split("=>")[1].split()[0]) == os.path.realpath(
        os.path.join(ROOT, "cudajpeg_amd", "libcudajpeg_amd.so"))
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "USAGE" in out.stderr, (out.returncode, out.stderr)
