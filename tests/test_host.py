"""Host-side logic of the product library on the CPU: the C ABI loads and
exports every declared symbol, the parser agrees with the oracle and reports
the reference's error codes, and the synthetic encoder is deterministic."""
import ctypes
import io
import os
import re

import numpy as np
import pytest

import oracle
from cudajpeg_amd import _lib, parse, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared("cudajpeg_amd.h") | _declared("format.h")
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == names


def _frame(data):
    e, f = parse(data)
    return e, f


@pytest.mark.parametrize("spec", [
    synth.ImageSpec(64, 48, "420", 75, 6.0, 1),
    synth.ImageSpec(33, 17, "444", 95, 8.0, 2),
    synth.ImageSpec(40, 24, "gray", 50, 8.0, 3),
    synth.ImageSpec(1918, 1078, "420", 85, 8.0, 4),
    synth.ImageSpec(100, 60, "420", 70, 6.0, 9, restart_interval=3),
])
def test_parse_matches_oracle(spec):
    enc = synth.encode(spec)
    e, f = _frame(enc.data)
    p = oracle.parse(enc.data)
    assert e == 0 and p.error == 0
    assert (f.width, f.height, f.num_channels) == (p.width, p.height, p.num_channels)
    assert (f.mcus_x, f.mcus_y, f.num_blocks, f.restart_interval) == (p.mcus_x, p.mcus_y, p.num_blocks,
                                                                      p.restart_interval)
    for c, ch in enumerate(p.channels):
        assert f.comp_stride[c] == ch["stride"] and f.comp_plane_height[c] == ch["plane_height"]
        assert f.comp_width[c] == ch["width"] and f.comp_height[c] == ch["height"]
        assert (f.dc_id[c], f.ac_id[c]) == (ch["dc_id"], ch["ac_id"])
    scan = enc.data[f.scan_offset: f.scan_offset + f.scan_length]
    assert scan == oracle.scan_bytes(p)


def _segments(data):
    out, i = [], 2
    while i < len(data) - 2:
        m = data[i + 1]
        L = (data[i + 2] << 8) | data[i + 3]
        out.append((i, m, L))
        if m == 0xDA:
            break
        i += 2 + L
    return out


def _patch(data, marker, off, value):
    b = bytearray(data)
    for i, m, L in _segments(data):
        if m == marker:
            b[i + 4 + off] = value
            return bytes(b)
    raise AssertionError(marker)


BASE = synth.encode(synth.ImageSpec(32, 32, "420", 75, 6.0, 5)).data


@pytest.mark.parametrize("name,data,code", [
    ("bad magic", b"\x00" + BASE[1:], 5),                         # src/format.cu:160-163
    ("missing EOI", BASE[:-2] + b"\x00\x00", 5),
    ("12-bit precision", _patch(BASE, 0xC0, 0, 12), 2),            # src/format.cu:248-249
    ("2 components", _patch(BASE, 0xC0, 5, 2), 2),                 # src/format.cu:257-258
    ("sampling 3x1", _patch(BASE, 0xC0, 7, 0x31), 2),               # src/format.cu:274-276
    ("dq id 4", _patch(BASE, 0xC0, 8, 4), 1),                      # src/format.cu:271-272
    ("16-bit DQT", _patch(BASE, 0xDB, 0, 0x10), 1),                 # src/format.cu:432
    ("DHT class 2", _patch(BASE, 0xC4, 0, 0x20), 1),                # src/format.cu:366
    ("DHT id 2", _patch(BASE, 0xC4, 0, 0x02), 2),                   # src/format.cu:367
    ("progressive SOF2", BASE.replace(b"\xff\xc0", b"\xff\xc2", 1), 1),  # src/format.cu:198-200
    ("Se != 63", _patch(BASE, 0xDA, 1 + 2 * 3 + 1, 62), 2),         # src/decodeScanGPU.cu:140
    ("scan component count", _patch(BASE, 0xDA, 0, 1), 2),          # src/decodeScanGPU.cu:131
    ("scan table id 2", _patch(BASE, 0xDA, 2, 0x22), 1),            # src/decodeScanGPU.cu:136
])
def test_parse_error_codes_match_reference(name, data, code):
    e, _ = _frame(data)
    assert e == code, name
    # the oracle restates the same checks
    assert oracle.parse(data).error == code, name


def test_synth_deterministic_and_decodable():
    from PIL import Image
    s = synth.ImageSpec(96, 80, "420", 80, 6.0, 42)
    a, b = synth.encode(s), synth.encode(s)
    assert a.data == b.data
    assert synth.encode(synth.ImageSpec(96, 80, "420", 80, 6.0, 43)).data != a.data
    im = Image.open(io.BytesIO(a.data))
    assert im.size == (96, 80)
    q = synth.quant_table(75, False)
    assert q[0, 0] == 8 and q[7, 7] == 50   # IJG scaling of Annex K at q75


def test_mixed_distribution_shape():
    specs = synth.config_mixed(256)
    assert len(specs) == 256
    areas = np.array([s.width * s.height for s in specs])
    assert areas.min() >= 200 * 200 and areas.max() <= 4200 * 3200
    frac420 = np.mean([s.sampling == "420" for s in specs])
    assert 0.55 < frac420 < 0.85


def _oversized_dht(data):
    """`data` with its first DHT replaced by one whose Kraft sum holds (255
    codes at each length 9..16) but which has 2040 symbols, more than the 256
    a HUFFVAL table can hold."""
    for i, m, L in _segments(data):
        if m == 0xC4:
            body = bytes([0x00]) + bytes([0] * 8 + [255] * 8) + (bytes(range(256)) * 8)[:2040]
            seg = b"\xff\xc4" + (2 + len(body)).to_bytes(2, "big") + body
            return data[:i] + seg + data[i + 2 + L:]
    raise AssertionError("no DHT")


def test_dht_with_more_than_256_symbols_is_a_syntax_error():
    bad = _oversized_dht(BASE)
    e, _ = _frame(bad)
    assert e == 1
    assert oracle.parse(bad).error == 1


def test_parser_under_address_sanitizer(tmp_path):
    """parse.cpp built for the host with AddressSanitizer parses the
    oversized DHT, the malformed headers above and a valid file without an
    out-of-bounds access."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("no host C++ compiler")
    drv = tmp_path / "drv.cpp"
    drv.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include "parse_internal.h"
int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    FILE* f = fopen(argv[a], "rb");
    std::vector<unsigned char> d;
    int c;
    while ((c = fgetc(f)) != EOF) d.push_back((unsigned char)c);
    fclose(f);
    cjd::ParseState st;
    cjd_frame_info fi;
    printf("%d\n", cjd::parse_until_scan(st, d.data(), d.size(), &fi));
  }
  return 0;
}
''')
    csrc = os.path.join(ROOT, "cudajpeg_amd", "csrc")
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-I" + csrc, str(drv), os.path.join(csrc, "parse.cpp"), "-o", str(exe)], check=True)
    cases = [(_oversized_dht(BASE), 1), (BASE, 0), (BASE[:-2] + b"\x00\x00", 5),
             (_patch(BASE, 0xC4, 0, 0x20), 1), (BASE[:200], None)]
    paths = []
    for k, (d, _) in enumerate(cases):
        p = tmp_path / f"c{k}.jpg"
        p.write_bytes(d)
        paths.append(str(p))
    r = subprocess.run([str(exe)] + paths, capture_output=True, text=True,
                       env={**os.environ, "ASAN_OPTIONS": "detect_leaks=0"})
    assert r.returncode == 0, r.stderr[-2000:]
    codes = [int(x) for x in r.stdout.split()]
    assert len(codes) == len(cases)
    for (d, want), got in zip(cases, codes):
        if want is not None:
            assert got == want


def test_dqt16_extension_in_host_parser():
    # the reference refuses 16-bit quant tables (src/format.cu:432); with
    # CJD_EXT_DQT16 the parser takes them: dq_tables16 holds the values,
    # dq_tables the 8-bit view clamped at 255
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle import _dqt16_tables
    from cudajpeg_amd import _lib as L
    from cudajpeg_amd.decoder import parse
    enc = synth.encode(synth.ImageSpec(120, 72, "420", 4, 6.0, 83, dqt16=1))
    e, _ = parse(enc.data)
    assert e == 1
    e, f = parse(enc.data, L.EXT_DQT16)
    assert e == 0
    for tq, t in _dqt16_tables(enc.data).items():
        assert list(f.dq_tables16[tq]) == t
        assert list(f.dq_tables[tq]) == [min(v, 255) for v in t]
    # an 8-bit file parses the same either way, dq_tables16 widened
    enc8 = synth.encode(synth.ImageSpec(64, 48, "420", 75, 6.0, 84))
    e0, f0 = parse(enc8.data)
    e1, f1 = parse(enc8.data, L.EXT_DQT16)
    assert e0 == e1 == 0
    assert [list(t) for t in f0.dq_tables] == [list(t) for t in f1.dq_tables]
    assert [list(t) for t in f1.dq_tables16] == [list(t) for t in f1.dq_tables]
