"""GPU test of the two-part decode (cjd_batch_set_overlap): the pixel kernel
of the first images runs on a second stream beside the entropy decode of the
rest.  Frames and coefficients must equal those of the two kernels run one
after the other, decode after decode."""
import numpy as np
import pytest

from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def files():
    # a mixed batch large enough to be split (at least 64 work items)
    return [e.data for e in synth.encode_many(synth.config_mixed(24, 5150))]


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_MFMA])
def test_two_part_decode_equals_one_after_the_other(gpu, files, idct):
    got = []
    for on in (False, True):
        b = BatchDecoder(idct, tile_bits=8192)
        b.set_tuning(0, 128)
        b.set_overlap(on)
        assert not b.stage(files).any()
        for _ in range(3):
            b.decode()
        errs = b.finish()
        print("overlap", on, "errors", errs.tolist(), "fallbacks", b.fallbacks(), flush=True)
        assert not errs.any()
        got.append(([b.pixels(i) for i in range(len(files))], [b.coefficients(i) for i in range(len(files))]))
        b.close()
    (p0, c0), (p1, c1) = got
    for i in range(len(files)):
        assert np.array_equal(p0[i], p1[i]), i
        assert np.array_equal(c0[i], c1[i]), i
