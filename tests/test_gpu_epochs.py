"""GPU tests of the batch scratch that is never cleared between decodes.

Every entropy launch tags its lookback records and per-image status words
with an epoch of its own and takes its work-item tickets from a counter that
only grows, so a decode is the entropy and pixel launches alone.  These tests
decode the same batch many times in a row, re-stage a batch object with a
different number of images (so the status words and records move over what
the last batch left), and keep two batches in flight on two streams: errors
must be reported exactly where the oracle reports them and the frames must
stay bit-exact.
"""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_INT, BatchDecoder, synth

pytestmark = pytest.mark.gpu

SPECS = [
    synth.ImageSpec(64, 48, "420", 75, 6.0, 201),
    synth.ImageSpec(300, 200, "420", 95, 12.0, 202, random_tile_frac=0.3),
    synth.ImageSpec(37, 29, "422", 80, 5.0, 203),
    synth.ImageSpec(100, 60, "420", 70, 6.0, 204, restart_interval=3),
    synth.ImageSpec(320, 240, "444", 90, 8.0, 205, huff=1),
    synth.ImageSpec(40, 24, "gray", 50, 8.0, 206),
]


def _corrupt(seed):
    enc = synth.encode(synth.ImageSpec(96, 64, "420", 75, 6.0, seed))
    data = bytearray(enc.data)
    p = oracle.parse(bytes(data))
    start = len(data) - 2 - p._raw.scan_len
    mid = start + p._raw.scan_len // 2
    data[mid:mid + 6] = b"\xff\x00\xff\x00\xff\x00"
    return bytes(data), oracle.decode_scan(oracle.parse(bytes(data))).error


@pytest.fixture(scope="module")
def encoded():
    return [synth.encode(s, truth=True) for s in SPECS]


def _ref(enc):
    p = oracle.parse(enc.data)
    return oracle.colour(p, oracle.planes(p, enc.coefs, 0))


def test_repeated_decodes_report_only_their_own_errors(gpu, encoded):
    bad, e_bad = _corrupt(211)
    assert e_bad != 0
    files = [e.data for e in encoded[:3]] + [bad] + [e.data for e in encoded[3:]]
    want = np.array([0, 0, 0, e_bad, 0, 0, 0], np.int32)
    b = BatchDecoder(IDCT_INT)
    assert np.array_equal(b.stage(files), np.zeros(len(files), np.int32))
    for k in range(24):
        b.decode()
        if k % 3 == 2:
            got = b.finish()
            print("decode", k, "errors", got.tolist(), flush=True)
            assert np.array_equal(got, want), k
    got = b.finish()
    assert np.array_equal(got, want)
    for i, enc in zip([0, 1, 2, 4, 5, 6], encoded):
        assert np.array_equal(b.pixels(i), _ref(enc)), i
        assert np.array_equal(b.coefficients(i), enc.coefs), i
    b.close()


def test_restage_with_another_layout(gpu, encoded):
    # one batch object, staged in turn with 2, 7 and 1 images: each stage
    # moves the status words and lookback records over the last batch's
    bad, e_bad = _corrupt(212)
    good = [e.data for e in encoded]
    rounds = [
        ([bad, good[0]], [e_bad, 0]),
        (good + [good[1]], [0] * 7),
        ([bad], [e_bad]),
        (good[::-1], [0] * 6),
    ]
    b = BatchDecoder(IDCT_INT)
    for r, (files, want) in enumerate(rounds):
        print("stage", r, "of", len(files), "images", flush=True)
        assert not b.stage(files).any()
        for _ in range(3):
            b.decode()
        assert b.finish().tolist() == want, r
    for i, enc in enumerate(encoded[::-1]):
        assert np.array_equal(b.pixels(i), _ref(enc)), i
    b.close()


def test_two_batches_in_flight_on_two_streams(gpu):
    import torch
    specs = synth.config_mixed(12, 4321)
    files = [e.data for e in synth.encode_many(specs)]
    ref = BatchDecoder(IDCT_INT)
    assert not ref.stage(files).any()
    ref.decode()
    assert not ref.finish().any()
    want = [ref.pixels(i) for i in range(len(files))]
    ref.close()
    decs = [BatchDecoder(IDCT_INT), BatchDecoder(IDCT_INT)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for d in decs:
        assert not d.stage(files).any()
    for k in range(16):
        decs[k % 2].decode(streams[k % 2])
    for k, (d, s) in enumerate(zip(decs, streams)):
        got = d.finish(s)
        print("decoder", k, "errors", got.tolist(), "fallbacks", d.fallbacks(), flush=True)
        assert not got.any()
        for i in range(len(files)):
            assert np.array_equal(d.pixels(i), want[i]), (k, i)
        d.close()
