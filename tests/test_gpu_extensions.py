"""Decoding beyond the reference, behind CJD_EXT_* flags (off by default).

16-bit quantisation tables (DQT Pq = 1, refused by src/format.cu:432): with
CJD_EXT_DQT16 a batch decodes them -- coefficients equal the encoder's,
integer-iDCT pixels the oracle's (decoding with the same extension), FP32 and
MFMA planes within 1 of the oracle's FP64 planes; without it the same files
are refused with SYNTAX_ERROR exactly as the reference refuses them.  The
reader API takes the extension through JPGReader.extensions."""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_FLOAT, IDCT_INT, IDCT_MFMA, BatchDecoder, JPGReader, synth
from cudajpeg_amd._lib import EXT_DQT16

pytestmark = pytest.mark.gpu

SPECS = [synth.ImageSpec(96, 64, "gray", 3, 6.0, 91, dqt16=1),
         synth.ImageSpec(130, 70, "420", 5, 8.0, 92, dqt16=1, qtables=3),
         synth.ImageSpec(64, 64, "444", 8, 10.0, 93, dqt16=1, restart_interval=5),
         synth.ImageSpec(1920, 1080, "420", 4, 8.0, 94, dqt16=1),
         synth.ImageSpec(800, 600, "422", 75, 8.0, 95)]  # 8-bit tables in the same batch


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_FLOAT, IDCT_MFMA])
def test_dqt16_batch(gpu, idct):
    encs = [synth.encode(s, truth=True) for s in SPECS]
    b = BatchDecoder(idct, planes=idct != IDCT_INT)
    errs = b.stage([e.data for e in encs])
    assert list(errs) == [1, 1, 1, 1, 0]  # the reference's behaviour by default
    b.set_extensions(EXT_DQT16)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    for i, (s, e) in enumerate(zip(SPECS, encs)):
        print("case dqt16", idct, s, flush=True)
        p = oracle.parse(e.data, oracle.EXT_DQT16)
        assert np.array_equal(b.coefficients(i), e.coefs), s
        if idct == IDCT_INT:
            ref = oracle.colour(p, oracle.planes(p, e.coefs, 0))
            assert np.array_equal(b.pixels(i), ref), s
        else:
            f64 = oracle.planes(p, e.coefs, 1)
            for c in range(p.num_channels):
                d = np.abs(b.plane(i, c).astype(int) - f64[c])
                assert d.max() <= 1, (s, c)
    b.close()


def test_dqt16_reader_api(gpu):
    enc = synth.encode(SPECS[1], truth=True)
    r = JPGReader(IDCT_INT, keep_coefficients=True)
    assert r.open(enc.data) == 1
    r.close()
    r = JPGReader(IDCT_INT, keep_coefficients=True, extensions=EXT_DQT16)
    assert r.open(enc.data) == 0
    p = oracle.parse(enc.data, oracle.EXT_DQT16)
    assert np.array_equal(r.pixels, oracle.colour(p, oracle.planes(p, enc.coefs, 0)))
    work = oracle.decode_scan(p, want_work=True).work
    for c in range(p.num_channels):
        assert np.array_equal(r.working_space(c), work[c]), c
    r.close()
