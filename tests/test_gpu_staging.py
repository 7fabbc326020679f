"""Staging through pinned host buffers (cjd_batch_stage / _stage_async).

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
