"""The oracle (CPU restatement of the reference) pinned against the golden
fixtures written by the reference's own code, the encoder's ground truth and
closed forms.  CPU only."""
import os

import numpy as np
import pytest

import oracle
from cudajpeg_amd import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")


def _golden():
    g = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in g.files})
    return g, names


def test_golden_scan_decode_matches_reference_decodeScanCPU():
    # working_space (dequantised coefficients, reference layout) written by the
    # reference src/decodeScanCPU.cu
    g, names = _golden()
    assert len(names) >= 8
    for n in names:
        data = g[f"{n}/file"].tobytes()
        p = oracle.parse(data)
        assert p.error == 0, n
        r = oracle.decode_scan(p, want_work=True)
        assert r.error == 0, n
        for c in range(p.num_channels):
            assert np.array_equal(r.work[c], g[f"{n}/work{c}"]), (n, c)


def test_golden_pixels_match_reference_cpuonly():
    # RGB / gray frames written by the reference CPUonly decoder
    g, names = _golden()
    seen = 0
    for n in names:
        if f"{n}/pixels" not in g.files:
            continue
        data = g[f"{n}/file"].tobytes()
        e, pix = oracle.decode(data, 0)
        assert e == 0
        assert np.array_equal(pix, g[f"{n}/pixels"]), n
        seen += 1
    assert seen >= 7


@pytest.mark.parametrize("spec", [
    synth.ImageSpec(64, 48, "420", 75, 6.0, 1),
    synth.ImageSpec(33, 17, "444", 95, 8.0, 2),
    synth.ImageSpec(40, 24, "gray", 50, 8.0, 3),
    synth.ImageSpec(37, 29, "422", 80, 5.0, 7),
    synth.ImageSpec(100, 60, "420", 70, 6.0, 9, restart_interval=3),
    synth.ImageSpec(160, 120, "420", 95, 12.0, 11, random_tile_frac=0.3),
    synth.ImageSpec(512, 512, "420", 75, 6.0, 1000),
    synth.ImageSpec(131, 77, "411", 85, 6.0, 41),
    synth.ImageSpec(90, 101, "440", 75, 6.0, 42),
    synth.ImageSpec(150, 66, "410", 90, 8.0, 43, restart_interval=4),
])
def test_oracle_matches_encoder_truth(spec):
    enc = synth.encode(spec, truth=True)
    p = oracle.parse(enc.data)
    r = oracle.decode_scan(p)
    assert r.error == 0
    assert np.array_equal(r.qcoef, enc.coefs)
    assert np.array_equal(r.bitpos, enc.block_bits)


def test_idct_shortcut_is_exact():
    # CPUonly's all-zero-AC shortcuts (pixelTransformCPU.c:16-26, 62-76) equal
    # the full integer algorithm of the GPU tree (pixelTransformGPU.cu:25-142)
    rng = np.random.default_rng(5)
    for _ in range(2000):
        blk = np.zeros(64, np.int32)
        k = rng.integers(0, 6)
        idx = rng.choice(64, size=k, replace=False)
        blk[idx] = rng.integers(-600, 600, size=k)
        blk[0] = rng.integers(-1024, 1024)
        # rows with only the DC term exercise the shortcut
        if rng.uniform() < 0.5:
            blk[1:] = 0
        assert np.array_equal(oracle.idct(blk, "int"), oracle.idct(blk, "int_noshortcut"))


def test_f64_idct_matches_matrix_form():
    # T.81 A.3.3 written as C^T S C with numpy; a transposed operand or a
    # missing C(0) = 1/sqrt(2) fails this
    C = np.array([[(np.sqrt(0.125) if u == 0 else 0.5) * np.cos((2 * x + 1) * u * np.pi / 16)
                   for x in range(8)] for u in range(8)])
    rng = np.random.default_rng(6)
    for _ in range(200):
        S = rng.integers(-300, 300, size=(8, 8)).astype(np.int32)
        want = C.T @ S @ C + 128.0
        assert np.allclose(oracle.idct_f64_raw(S), want, atol=1e-9)
    # a DC-only block is flat at DC/8 + 128
    S = np.zeros((8, 8), np.int32)
    S[0, 0] = 80
    assert np.allclose(oracle.idct_f64_raw(S), 138.0)


def test_integer_idct_within_one_of_exact():
    enc = synth.encode(synth.ImageSpec(256, 192, "444", 90, 8.0, 77), truth=True)
    p = oracle.parse(enc.data)
    a = oracle.planes(p, enc.coefs, 0)
    b = oracle.planes(p, enc.coefs, 1)
    for c in range(3):
        d = np.abs(a[c].astype(int) - b[c])
        assert d.max() <= 1
        assert (d == 0).mean() > 0.97


def test_colour_closed_form():
    # grey chroma (128) leaves RGB = Y; pure Cr/Cb steps follow the reference gains
    enc = synth.encode(synth.ImageSpec(16, 16, "444", 90, 0.0, 3, amp_total=0.0), truth=True)
    p = oracle.parse(enc.data)
    pl = [np.full((16, 16), v, np.uint8) for v in (100, 128, 128)]
    rgb = oracle.colour(p, pl)
    assert (rgb == 100).all()
    pl = [np.full((16, 16), v, np.uint8) for v in (100, 128, 138)]
    rgb = oracle.colour(p, pl)
    assert rgb[0, 0, 0] == (100 * 256 + 359 * 10 + 128) >> 8
    assert rgb[0, 0, 1] == (100 * 256 - 183 * 10 + 128) >> 8
    assert rgb[0, 0, 2] == 100


def test_unstuff_rule():
    # src/decodeScanGPU.cu:11-24: FF00 -> FF, RSTn kept, other marker byte dropped
    assert oracle.unstuff(b"\x12\xff\x00\x34") == b"\x12\xff\x34"
    assert oracle.unstuff(b"\xff\xd3\x01") == b"\xff\xd3\x01"
    assert oracle.unstuff(b"\xff\xe1\x01") == b"\xff\x01"


def _dqt16_tables(data: bytes):
    """The file's quant tables read straight from its DQT segments (T.81
    B.2.4.1), independent of both parsers: {id: 64 values in zigzag order}."""
    out, pos = {}, 2
    while data[pos + 1] != 0xDA:
        n = (data[pos + 2] << 8) | data[pos + 3]
        if data[pos + 1] == 0xDB:
            q, end = pos + 4, pos + 2 + n
            while q < end:
                pq, tq = data[q] >> 4, data[q] & 15
                w = 2 if pq else 1
                out[tq] = [int.from_bytes(data[q + 1 + w * k:q + 1 + w * k + w], "big") for k in range(64)]
                q += 1 + 64 * w
        pos += 2 + n
    return out


def test_dqt16_extension_pinned():
    # 16-bit quant tables (Pq = 1): refused as the reference refuses them
    # (SYNTAX_ERROR, src/format.cu:432) unless the extension is asked for;
    # with it, the tables are the file's values and the FP64 planes equal
    # C^T (coef x q) C + 128 computed here from those values
    zz = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7,
          14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39,
          46, 53, 60, 61, 54, 47, 55, 62, 63]
    C = np.array([[(np.sqrt(0.125) if u == 0 else 0.5) * np.cos((2 * x + 1) * u * np.pi / 16)
                   for x in range(8)] for u in range(8)])
    for spec in [synth.ImageSpec(96, 64, "gray", 3, 6.0, 81, dqt16=1),
                 synth.ImageSpec(80, 48, "444", 5, 8.0, 82, dqt16=1, qtables=3)]:
        enc = synth.encode(spec, truth=True)
        assert oracle.parse(enc.data).error == 1
        p = oracle.parse(enc.data, oracle.EXT_DQT16)
        assert p.error == 0
        tabs = _dqt16_tables(enc.data)
        assert max(max(t) for t in tabs.values()) > 255  # really 16-bit content
        for tq, t in tabs.items():
            assert list(p.dq[tq]) == t
        f64 = oracle.planes(p, enc.coefs, 1)
        i32 = oracle.planes(p, enc.coefs, 0)
        # block 0 of component 0 by the closed form
        q = np.zeros(64)
        for k in range(64):
            q[zz[k]] = tabs[p.channels[0]["dq_id"]][k]
        S = enc.coefs[0].astype(np.float64) * q
        want = np.clip(np.floor(C.T @ S.reshape(8, 8) @ C + 128.0 + 0.5), 0, 255)
        assert np.array_equal(f64[0][:8, :8].astype(int), want.astype(int))
        for c in range(p.num_channels):
            assert np.abs(f64[c].astype(int) - i32[c]).max() <= 1
        err, px = oracle.decode(enc.data, 0, oracle.EXT_DQT16)
        assert err == 0 and px.shape == (spec.height, spec.width, spec.ncomp)
