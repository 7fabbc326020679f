"""Staging through pinned host buffers (cjd_batch_stage / _stage_async), and
device-side unstuffing against the host's.

A batch staged asynchronously on one stream while another batch decodes on a
second stream, then decoded on the first, gives exactly the frames and
coefficients of a batch staged synchronously; so does a batch re-staged with
larger files (its pinned buffers grow) and with the smaller ones again, with
no wait by the caller between the stages (the library waits for the last
stage's DMAs itself before it writes the pinned buffers again)."""
import numpy as np
import pytest
import torch

from cudajpeg_amd import IDCT_INT, BatchDecoder, synth
from cudajpeg_amd._lib import UNSTUFF_DEVICE, UNSTUFF_HOST

pytestmark = pytest.mark.gpu


def _frames(d, n):
    return [d.pixels(i) for i in range(n)], [d.coefficients(i) for i in range(n)]


def test_async_staging_equals_sync(gpu):
    small = synth.encode_many(synth.config_mixed(20, 8100), truth=True)
    large = synth.encode_many(synth.config_mixed(6, 8200)
                              + [synth.ImageSpec(4096, 3072, "420", 90, 9.0, 8201)], truth=True)
    busy = synth.encode_many(synth.config_mixed(16, 8300))
    refs = {}
    for name, encs in (("small", small), ("large", large)):
        r = BatchDecoder(IDCT_INT)
        assert not r.stage([e.data for e in encs]).any()
        r.decode()
        assert not r.finish().any()
        refs[name] = _frames(r, len(encs))
        for i, e in enumerate(encs):
            assert np.array_equal(refs[name][1][i], e.coefs), (name, i)
        r.close()
    other = BatchDecoder(IDCT_INT)
    assert not other.stage([e.data for e in busy]).any()
    d = BatchDecoder(IDCT_INT)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep, name in enumerate(["small", "large", "small", "large"]):
        encs = small if name == "small" else large
        print("case async staging rep", rep, name, flush=True)
        other.decode(s2)                                              # a decode running meanwhile
        assert not d.stage([e.data for e in encs], s1, asynchronous=True).any()
        d.decode(s1)                                                  # ordered after the uploads on s1
        assert not d.finish(s1).any()
        assert not other.finish(s2).any()
        px, cf = _frames(d, len(encs))
        for i in range(len(encs)):
            assert np.array_equal(cf[i], refs[name][1][i]), (rep, name, i)
            assert np.array_equal(px[i], refs[name][0][i]), (rep, name, i)
    # two stages in a row with no decode or wait between them
    assert not d.stage([e.data for e in large], s1, asynchronous=True).any()
    assert not d.stage([e.data for e in small], s1, asynchronous=True).any()
    d.decode(s1)
    assert not d.finish(s1).any()
    px, cf = _frames(d, len(small))
    for i in range(len(small)):
        assert np.array_equal(px[i], refs["small"][0][i]), i
    d.close()
    other.close()


def test_device_unstuff_equals_host(gpu):
    # the same batch staged with the scans unstuffed by the host planner and
    # by k_unstuff: identical coefficients and frames, restart intervals and
    # random tiles included (the words themselves are checked against the
    # host's in the CPU emulator of k_unstuff)
    specs = synth.config_mixed(40, 8400) + [
        synth.ImageSpec(1000, 700, "422", 80, 6.0, 8401, restart_interval=7),
        synth.ImageSpec(640, 480, "444", 95, 12.0, 8402, random_tile_frac=0.5),
        synth.ImageSpec(333, 211, "gray", 90, 8.0, 8403, restart_interval=1)]
    encs = synth.encode_many(specs, truth=True)
    out = {}
    for mode in (UNSTUFF_HOST, UNSTUFF_DEVICE):
        d = BatchDecoder(IDCT_INT)
        d.set_unstuff(mode)
        print("case unstuff mode", mode, flush=True)
        assert not d.stage([e.data for e in encs]).any()
        d.decode()
        assert not d.finish().any()
        out[mode] = _frames(d, len(encs))
        d.close()
    for i, e in enumerate(encs):
        assert np.array_equal(out[UNSTUFF_DEVICE][1][i], e.coefs), (i, specs[i])
        assert np.array_equal(out[UNSTUFF_DEVICE][1][i], out[UNSTUFF_HOST][1][i]), i
        assert np.array_equal(out[UNSTUFF_DEVICE][0][i], out[UNSTUFF_HOST][0][i]), i
