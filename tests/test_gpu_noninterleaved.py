"""Non-interleaved baseline scans on the GPU, behind CJD_EXT_NONINTERLEAVED.

One scan per component (T.81 A.2.2), which the reference refuses
(src/decodeScanGPU.cu:124-131).  The oracle's decode of such files is pinned
against the same images encoded as one interleaved scan, in
tests/test_noninterleaved.py.  Here the batch decodes them beside
interleaved images; without the flag they are refused with SYNTAX_ERROR, as
the reference refuses them.  With the flag:
* coefficients equal the encoder's (zero where MCUs pad a component's grid);
* integer-iDCT pixels equal the oracle's;
* FP32 and MFMA planes are within 1 of the oracle's FP64 planes.
Both entropy decoders, graph replays and the reader API are covered.  The
encoder redefines Huffman table 0, the restart interval and (second form)
quant table 1 between scans."""
import numpy as np
import pytest
import torch

import oracle
from cudajpeg_amd import ENTROPY_SERIAL, IDCT_FLOAT, IDCT_INT, IDCT_MFMA, BatchDecoder, JPGReader, synth
from cudajpeg_amd._lib import EXT_NONINTERLEAVED

pytestmark = pytest.mark.gpu

NI = oracle.EXT_NONINTERLEAVED

SPECS = [synth.ImageSpec(333, 217, "420", 80, 6.0, 41, noninterleaved=1),
         synth.ImageSpec(17, 9, "422", 75, 8.0, 42, noninterleaved=2),
         synth.ImageSpec(640, 480, "444", 90, 10.0, 43, restart_interval=7, noninterleaved=2),
         synth.ImageSpec(1001, 77, "411", 70, 6.0, 44, restart_interval=3, noninterleaved=1),
         synth.ImageSpec(300, 500, "440", 85, 6.0, 45, noninterleaved=1),
         synth.ImageSpec(260, 130, "410", 60, 12.0, 46, restart_interval=5, noninterleaved=2),
         synth.ImageSpec(800, 600, "420", 75, 8.0, 47),  # interleaved files in the same batch
         synth.ImageSpec(96, 64, "gray", 50, 6.0, 48)]
IS_NI = [s.noninterleaved > 0 for s in SPECS]


def _check(b, i, s, e, idct):
    p = oracle.parse(e.data, NI)
    assert p.error == 0, s
    assert np.array_equal(b.coefficients(i), e.coefs), s
    if idct == IDCT_INT:
        assert np.array_equal(b.pixels(i), oracle.colour(p, oracle.planes(p, e.coefs, 0))), s
    else:
        f64 = oracle.planes(p, e.coefs, 1)
        for c in range(p.num_channels):
            d = np.abs(b.plane(i, c).astype(int) - f64[c])
            assert d.max() <= 1, (s, c)


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_FLOAT, IDCT_MFMA])
def test_noninterleaved_batch(gpu, idct):
    encs = [synth.encode(s, truth=True) for s in SPECS]
    b = BatchDecoder(idct, tile_bits=512, planes=idct != IDCT_INT)
    errs = b.stage([e.data for e in encs])
    assert list(errs) == [1 if n else 0 for n in IS_NI]  # the reference's behaviour by default
    b.set_extensions(EXT_NONINTERLEAVED)
    assert not b.stage([e.data for e in encs]).any()
    for rep in range(2):  # the second decode reuses the staged records
        b.decode()
        assert not b.finish().any()
        for i, (s, e) in enumerate(zip(SPECS, encs)):
            print("case noninterleaved", idct, rep, s, flush=True)
            _check(b, i, s, e, idct)
    b.close()


def test_noninterleaved_large_multi_item(gpu):
    """A 4096x3072 4:2:0 image in 2048-bit tiles: every scan segment spans
    many work items (lookback across items), with and without restarts."""
    specs = [synth.ImageSpec(4096, 3072, "420", 85, 8.0, 61, noninterleaved=2),
             synth.ImageSpec(4000, 3000, "444", 90, 6.0, 62, restart_interval=200, noninterleaved=1)]
    encs = [synth.encode(s, truth=True) for s in specs]
    b = BatchDecoder(IDCT_INT, tile_bits=2048)
    b.set_extensions(EXT_NONINTERLEAVED)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    for i, (s, e) in enumerate(zip(specs, encs)):
        print("case noninterleaved large", s, flush=True)
        _check(b, i, s, e, IDCT_INT)
    b.close()


def test_noninterleaved_serial_entropy(gpu):
    encs = [synth.encode(s, truth=True) for s in SPECS]
    b = BatchDecoder(IDCT_INT, entropy=ENTROPY_SERIAL)
    b.set_extensions(EXT_NONINTERLEAVED)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    for i, (s, e) in enumerate(zip(SPECS, encs)):
        print("case noninterleaved serial", s, flush=True)
        _check(b, i, s, e, IDCT_INT)
    b.close()


def test_noninterleaved_graph_replays(gpu):
    encs = [synth.encode(s, truth=True) for s in SPECS]
    b = BatchDecoder(IDCT_INT, tile_bits=1024)
    b.set_extensions(EXT_NONINTERLEAVED)
    assert not b.stage([e.data for e in encs]).any()
    st = torch.cuda.Stream()
    b.decode(st)
    assert not b.finish(st).any()
    b.capture(st)
    for rep in range(3):
        b.replay(st)
        assert not b.finish(st).any()
        torch.cuda.synchronize()
        for i, (s, e) in enumerate(zip(SPECS, encs)):
            print("case noninterleaved replay", rep, s, flush=True)
            _check(b, i, s, e, IDCT_INT)
    b.close()


def test_noninterleaved_reader_api(gpu):
    enc = synth.encode(SPECS[2], truth=True)
    r = JPGReader(IDCT_INT, keep_coefficients=True)
    assert r.open(enc.data) == 1
    r.close()
    r = JPGReader(IDCT_INT, keep_coefficients=True, extensions=EXT_NONINTERLEAVED)
    assert r.open(enc.data) == 0
    p = oracle.parse(enc.data, NI)
    assert np.array_equal(r.pixels, oracle.colour(p, oracle.planes(p, enc.coefs, 0)))
    work = oracle.decode_scan(p, want_work=True).work
    for c in range(p.num_channels):
        assert np.array_equal(r.working_space(c), work[c]), c
    r.close()
