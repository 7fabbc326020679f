"""GPU parity of the HIP decoder against the oracle and the encoder's ground truth.

Coefficients and block order are bit-exact; integer-iDCT pixels are bit-exact
with the reference's integer Chen-Wang iDCT + colour transform; FP32-iDCT planar
samples are within 1 of it (the reference has no float iDCT, see the tolerance
note in the float test).
"""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import (ENTROPY_SERIAL, ENTROPY_TILED, IDCT_FLOAT, IDCT_INT, IDCT_MFMA, BatchDecoder, JPGReader,
                          synth)

pytestmark = pytest.mark.gpu

SMALL = [
    synth.ImageSpec(64, 48, "420", 75, 6.0, 1),
    synth.ImageSpec(33, 17, "444", 95, 8.0, 2),
    synth.ImageSpec(40, 24, "gray", 50, 8.0, 3),
    synth.ImageSpec(37, 29, "422", 80, 5.0, 7),
    synth.ImageSpec(100, 60, "420", 70, 6.0, 9, restart_interval=3),
    synth.ImageSpec(300, 200, "420", 95, 12.0, 11, random_tile_frac=0.3),
    synth.ImageSpec(17, 9, "420", 90, 4.0, 12),
    synth.ImageSpec(200, 136, "420", 80, 6.0, 31, qtables=3),
    synth.ImageSpec(72, 40, "444", 60, 6.0, 32, qtables=3),
    synth.ImageSpec(8, 8, "gray", 100, 2.0, 13),
    # long-code AC tables: more second-level lookup tables than the pool holds
    synth.ImageSpec(320, 240, "444", 90, 8.0, 71, huff=1),
    synth.ImageSpec(400, 300, "420", 75, 8.0, 72, huff=1, restart_interval=5),
]


def _ref_pixels(data, coefs):
    p = oracle.parse(data)
    return oracle.colour(p, oracle.planes(p, coefs, 0)), p


@pytest.fixture(scope="module")
def small_encoded():
    return [synth.encode(s, truth=True) for s in SMALL]


@pytest.mark.parametrize("entropy", [ENTROPY_TILED, ENTROPY_SERIAL])
def test_each_image_alone(gpu, small_encoded, entropy):
    for spec, enc in zip(SMALL, small_encoded):
        print("case", spec, flush=True)
        b = BatchDecoder(IDCT_INT, entropy, planes=True)
        assert b.stage([enc.data])[0] == 0
        b.decode()
        assert b.finish()[0] == 0
        assert np.array_equal(b.coefficients(0), enc.coefs), spec
        ref, p = _ref_pixels(enc.data, enc.coefs)
        assert np.array_equal(b.pixels(0), ref), spec
        planes = oracle.planes(p, enc.coefs, 0)
        for c in range(p.num_channels):
            assert np.array_equal(b.plane(0, c), planes[c]), (spec, c)
        b.close()


def test_batch_equals_alone(gpu, small_encoded):
    b = BatchDecoder(IDCT_INT)
    assert not b.stage([e.data for e in small_encoded]).any()
    b.decode()
    assert not b.finish().any()
    for i, (spec, enc) in enumerate(zip(SMALL, small_encoded)):
        assert np.array_equal(b.coefficients(i), enc.coefs), spec
        ref, _ = _ref_pixels(enc.data, enc.coefs)
        assert np.array_equal(b.pixels(i), ref), spec
    assert b.fallbacks() == 0


@pytest.mark.parametrize("tile_bits", [2048, 4096, 8192])
def test_large_configs(gpu, tile_bits):
    specs = [synth.config_plain_512(), synth.config_4k()] + synth.config_shapes()
    encs = synth.encode_many(specs, truth=True)
    b = BatchDecoder(IDCT_INT, ENTROPY_TILED, tile_bits)
    for spec, enc in zip(specs, encs):
        print("case", spec.width, spec.height, spec.sampling, "tile", tile_bits, flush=True)
        assert b.stage([enc.data])[0] == 0
        b.decode()
        assert b.finish()[0] == 0
        assert np.array_equal(b.coefficients(0), enc.coefs)
        ref, _ = _ref_pixels(enc.data, enc.coefs)
        assert np.array_equal(b.pixels(0), ref)
    assert b.fallbacks() == 0


@pytest.mark.parametrize("idct", [IDCT_FLOAT, IDCT_MFMA])
def test_float_idct_within_one(gpu, small_encoded, idct):
    # The reference's iDCT is integer (CPUonly/src/pixelTransformCPU.c:14-107);
    # an exact iDCT rounded to nearest differs from it by at most 1 per planar
    # sample.  Tolerance: planar |d| <= 1 against the integer iDCT and against
    # the FP64 iDCT; RGB is reported, not asserted beyond the colour gains
    # (359/256, 454/256 amplify a 1-LSB chroma step to <= 3 LSB).
    for spec, enc in zip(SMALL, small_encoded):
        b = BatchDecoder(idct, planes=True)
        assert b.stage([enc.data])[0] == 0
        b.decode()
        assert b.finish()[0] == 0
        p = oracle.parse(enc.data)
        pl_int = oracle.planes(p, enc.coefs, 0)
        pl_f64 = oracle.planes(p, enc.coefs, 1)
        for c in range(p.num_channels):
            got = b.plane(0, c).astype(int)
            assert np.abs(got - pl_int[c]).max() <= 1, spec
            assert np.abs(got - pl_f64[c]).max() <= 1, spec
        ref, _ = _ref_pixels(enc.data, enc.coefs)
        assert np.abs(b.pixels(0).astype(int) - ref).max() <= 3
        b.close()


def test_reader_api(gpu, small_encoded):
    r = JPGReader(IDCT_INT, keep_coefficients=True)
    for spec, enc in zip(SMALL, small_encoded):
        assert r.open(enc.data) == 0, (spec, r.error)
        ref, p = _ref_pixels(enc.data, enc.coefs)
        assert np.array_equal(r.pixels, ref), spec
        work = oracle.decode_scan(p, want_work=True).work
        for c in range(p.num_channels):
            assert np.array_equal(r.working_space(c), work[c]), (spec, c)
    r.close()


def test_corrupt_scan_reports_syntax(gpu):
    enc = synth.encode(synth.ImageSpec(64, 64, "420", 75, 6.0, 21))
    data = bytearray(enc.data)
    p = oracle.parse(bytes(data))
    # overwrite the middle of the scan with bytes that contain an invalid code
    start = len(data) - 2 - p._raw.scan_len
    mid = start + p._raw.scan_len // 2
    data[mid:mid + 6] = b"\xff\x00\xff\x00\xff\x00"
    e_ref = oracle.decode_scan(oracle.parse(bytes(data))).error
    b = BatchDecoder(IDCT_INT)
    assert b.stage([bytes(data)])[0] == 0
    b.decode()
    assert b.finish()[0] == e_ref


def test_mfma_idct_error_vs_fp64(gpu):
    # FP16 hi/lo basis with FP32 accumulation: planar samples within 1 of the
    # FP64 iDCT (T.81 A.3.3) and >= 99.9% equal to it; within 1 of the
    # reference's integer iDCT
    specs = [synth.config_plain_512(), synth.config_4k(), synth.ImageSpec(640, 480, "444", 95, 8.0, 5),
             synth.ImageSpec(512, 384, "gray", 60, 8.0, 6),
             synth.ImageSpec(800, 600, "420", 85, 8.0, 8, qtables=3)]
    for spec in specs:
        enc = synth.encode(spec, truth=True)
        b = BatchDecoder(IDCT_MFMA, planes=True)
        assert b.stage([enc.data])[0] == 0
        b.decode()
        assert b.finish()[0] == 0
        p = oracle.parse(enc.data)
        f64 = oracle.planes(p, enc.coefs, 1)
        i32 = oracle.planes(p, enc.coefs, 0)
        for c in range(p.num_channels):
            got = b.plane(0, c).astype(int)
            d = np.abs(got - f64[c])
            print(spec.width, spec.height, spec.sampling, c, "max", d.max(), "exact", (d == 0).mean(), flush=True)
            assert d.max() <= 1 and (d == 0).mean() > 0.999
            assert np.abs(got - i32[c]).max() <= 1
        b.close()


def test_tune_table_per_image_parameters(gpu):
    # a per-scan-size table puts the small and the large images of one batch
    # on different tile lengths / item sizes; coefficients stay bit-exact
    specs = [synth.ImageSpec(96, 64, "420", 75, 6.0, 61), synth.ImageSpec(640, 480, "444", 90, 8.0, 62),
             synth.ImageSpec(1024, 768, "420", 85, 10.0, 63, random_tile_frac=0.2)]
    encs = [synth.encode(s, truth=True) for s in specs]
    table = [(8 << 10, 1024, 0, 16), (256 << 10, 2048, 256, 64), (1 << 62, 8192, 0, 256)]
    b = BatchDecoder(IDCT_INT)
    b.set_tune_table(table)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    for i, (s, e) in enumerate(zip(specs, encs)):
        assert np.array_equal(b.coefficients(i), e.coefs), s
        ref, _ = _ref_pixels(e.data, e.coefs)
        assert np.array_equal(b.pixels(i), ref), s
    b.close()
