"""The allocation-free entry point cjd_decode_noalloc against the batch entry
points (cjd_batch_stage + decode + finish) on the same files: two calls in a
row, the second with a smaller image, each bit-identical to the batch path's
frame (integer and MFMA iDCT; the integer frames also equal the oracle's),
and a file whose decode needs more scratch than the array holds is refused
with OOM_ERROR before any device work."""
import ctypes

import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, _lib, synth

pytestmark = pytest.mark.gpu


def _noalloc(data, idct, cap_extra=0):
    import torch
    p = oracle.parse(data)
    ch = 3 if p.num_channels == 3 else 1
    frame = torch.full((p.height * p.width * ch + cap_extra,), 0xA5, dtype=torch.uint8, device="cuda")
    need = ctypes.c_size_t()
    stream = torch.cuda.current_stream()
    e = _lib.lib().cjd_decode_noalloc(0, data, len(data), idct, ctypes.c_void_p(frame.data_ptr()), frame.numel(),
                                      ctypes.c_void_p(stream.cuda_stream), ctypes.byref(need))
    out = frame[: p.height * p.width * ch].cpu().numpy().reshape(p.height, p.width, ch) if e == 0 else None
    return e, out, need.value


def _batch(data, idct):
    b = BatchDecoder(idct)
    assert b.stage([data])[0] == 0
    b.decode()
    assert b.finish()[0] == 0
    out = b.pixels(0)
    b.close()
    return out


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_MFMA])
def test_noalloc_equals_batch_path_two_calls(gpu, idct):
    specs = [synth.ImageSpec(1918, 1078, "420", 85, 8.0, 3002), synth.ImageSpec(512, 384, "444", 90, 6.0, 77)]
    cap = _lib.lib().cjd_noalloc_scratch_bytes()
    for s in specs:  # the second call decodes the smaller image over the first one's scratch
        print("case noalloc", s, "idct", idct, flush=True)
        e = synth.encode(s, truth=True)
        err, got, need = _noalloc(e.data, idct, cap_extra=4096)
        assert err == 0, err
        assert 0 < need <= cap
        assert np.array_equal(got, _batch(e.data, idct)), s
        if idct == IDCT_INT:
            p = oracle.parse(e.data)
            assert np.array_equal(got, oracle.colour(p, oracle.planes(p, e.coefs, 0))), s


def test_noalloc_refuses_oversize_and_bad_input(gpu):
    big = synth.encode(synth.ImageSpec(4096, 4096, "444", 95, 10.0, 78)).data
    print("case noalloc oversize", flush=True)
    err, _, need = _noalloc(big, IDCT_INT)
    assert err == 3 and need > _lib.lib().cjd_noalloc_scratch_bytes()
    # a frame buffer smaller than the frame is a caller error
    small = synth.encode(synth.ImageSpec(64, 48, "420", 75, 6.0, 1)).data
    err, _, _ = _noalloc(small, IDCT_INT, cap_extra=-1)
    assert err == -1
    # a file without a frame header gives the parser's code
    assert _noalloc(b"\xff\xd8\xff\xd9", IDCT_INT)[0] in (1, 5)
    # and the entry point still decodes afterwards
    e = synth.encode(synth.ImageSpec(64, 48, "420", 75, 6.0, 1), truth=True)
    err, got, _ = _noalloc(e.data, IDCT_INT)
    p = oracle.parse(e.data)
    assert err == 0 and np.array_equal(got, oracle.colour(p, oracle.planes(p, e.coefs, 0)))
