"""GPU parity on the power-of-two luma samplings beyond 4:2:0 / 4:2:2 / 4:4:4
that the reference accepts (src/format.cu:274-276): 4x1, 1x2 and 4x2 (6, 4
and 10 blocks per MCU), alone and in one batch with the common samplings.
Coefficients bit-exact with the encoder's ground truth; integer-iDCT pixels
bit-exact with the oracle, whose pixels on such files equal the reference
CPUonly decoder's (tests/golden)."""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import ENTROPY_SERIAL, ENTROPY_TILED, IDCT_INT, BatchDecoder, synth

pytestmark = pytest.mark.gpu

SPECS = [
    synth.ImageSpec(131, 77, "411", 85, 6.0, 41),
    synth.ImageSpec(90, 101, "440", 75, 6.0, 42),
    synth.ImageSpec(150, 66, "410", 90, 8.0, 43, restart_interval=4),
    synth.ImageSpec(37, 29, "411", 95, 8.0, 44),
    synth.ImageSpec(640, 360, "411", 80, 8.0, 45, random_tile_frac=0.2),
]


@pytest.fixture(scope="module")
def encoded():
    return [synth.encode(s, truth=True) for s in SPECS]


def _ref(enc):
    p = oracle.parse(enc.data)
    return oracle.colour(p, oracle.planes(p, enc.coefs, 0))


@pytest.mark.parametrize("entropy", [ENTROPY_TILED, ENTROPY_SERIAL])
def test_wide_sampling_alone(gpu, encoded, entropy):
    for spec, enc in zip(SPECS, encoded):
        print("case", spec, "entropy", entropy, flush=True)
        b = BatchDecoder(IDCT_INT, entropy)
        assert b.stage([enc.data])[0] == 0
        b.decode()
        assert b.finish()[0] == 0
        assert np.array_equal(b.coefficients(0), enc.coefs), spec
        assert np.array_equal(b.pixels(0), _ref(enc)), spec
        b.close()


def test_wide_sampling_in_a_mixed_batch(gpu, encoded):
    extra = [synth.ImageSpec(64, 48, "420", 75, 6.0, 1), synth.ImageSpec(40, 24, "gray", 50, 8.0, 3)]
    encs = list(encoded) + [synth.encode(s, truth=True) for s in extra]
    b = BatchDecoder(IDCT_INT, tile_bits=1024)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    for i, e in enumerate(encs):
        print("image", i, flush=True)
        assert np.array_equal(b.coefficients(i), e.coefs), i
        assert np.array_equal(b.pixels(i), _ref(e)), i
    b.close()
