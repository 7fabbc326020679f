"""GPU test of the decode in parts (cjd_batch_set_overlap): the pixel kernel
of each run of images but the last runs on a second stream beside the next
runs' entropy decodes.  Frames and coefficients must equal those of the two
kernels run one after the other, decode after decode."""
import numpy as np
import pytest

from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def files():
    # a mixed batch large enough to be split (at least 64 work items)
    return [e.data for e in synth.encode_many(synth.config_mixed(24, 5150))]


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_MFMA])
def test_decode_in_parts_equals_one_after_the_other(gpu, files, idct):
    got = []
    for parts in (1, 2, 3, 4):
        b = BatchDecoder(idct, tile_bits=8192)
        b.set_tuning(0, 128)
        assert not b.stage(files).any()
        b.set_overlap(parts)
        for _ in range(3):
            b.decode()
        errs = b.finish()
        print("parts", parts, "errors", errs.tolist(), "fallbacks", b.fallbacks(), flush=True)
        assert not errs.any()
        got.append(([b.pixels(i) for i in range(len(files))], [b.coefficients(i) for i in range(len(files))]))
        b.close()
    p0, c0 = got[0]
    for parts, (p1, c1) in zip((2, 3, 4), got[1:]):
        for i in range(len(files)):
            assert np.array_equal(p0[i], p1[i]), (parts, i)
            assert np.array_equal(c0[i], c1[i]), (parts, i)


def test_overlap_parts_out_of_range_is_refused(gpu):
    from cudajpeg_amd import DecodeError
    b = BatchDecoder()
    with pytest.raises(DecodeError):
        b.set_overlap(5)
    b.close()
