"""GPU tests of the lookback's bounded waits, of a multi-image MFMA batch, and
of the output handoff (PPM/PGM writer, zero-copy device views)."""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, JPGReader, device_frame, synth

pytestmark = pytest.mark.gpu

HARD = [synth.ImageSpec(1200, 900, "420", 63, 8.9, 5),
        synth.ImageSpec(640, 480, "444", 95, 12.0, 15, random_tile_frac=0.5),
        synth.ImageSpec(800, 600, "gray", 90, 12.0, 3),
        synth.ImageSpec(1000, 700, "422", 80, 6.0, 9, restart_interval=7)]


def _ref(enc):
    p = oracle.parse(enc.data)
    return oracle.colour(p, oracle.planes(p, enc.coefs, 0))


@pytest.mark.parametrize("wait_us", [0, 200000])
def test_lookback_wait_limit_does_not_change_results(gpu, wait_us):
    # Work items of 4 tiles of 2048 bits: many items per segment, so items
    # wait for their predecessors in the lookback.  With a limit of 0 an item
    # gives up whenever its predecessor has not published yet (forced from
    # the host); its image is then decoded again, alone, and then in ordered
    # launches, where no item waits: that redo resolves every image, so the
    # serial decoder is never needed.  Results are bit-exact either way and
    # no work item goes undecoded.
    encs = [synth.encode(s, truth=True) for s in HARD]
    b = BatchDecoder(IDCT_INT, tile_bits=2048)
    b.set_tuning(0, 4)
    b.set_wait_limit(wait_us)
    assert not b.stage([e.data for e in encs]).any()
    for rep in range(3):
        print("case wait_us", wait_us, "rep", rep, flush=True)
        b.decode()
        assert not b.finish().any()
        for i, (s, e) in enumerate(zip(HARD, encs)):
            assert np.array_equal(b.coefficients(i), e.coefs), s
            assert np.array_equal(b.pixels(i), _ref(e)), s
            assert not b.status(i) & 0x100, (s, hex(b.status(i)))  # no serial redo
        print("status", [hex(b.status(i)) for i in range(len(HARD))], flush=True)
    print("fallbacks", b.fallbacks(), flush=True)
    assert b.protocol_redos() == 0
    if wait_us:
        assert b.fallbacks() == 0
    b.close()


def test_mfma_multi_image_batch(gpu):
    # one MFMA batch: a file that fails to parse first, then images whose
    # per-image table operands sit at index > 0 -- three quant tables, gray,
    # 4:4:4 -- each within 1 of the FP64 and of the integer iDCT
    specs = [synth.ImageSpec(200, 136, "420", 80, 6.0, 31, qtables=3), synth.ImageSpec(512, 384, "gray", 60, 8.0, 6),
             synth.ImageSpec(640, 480, "444", 95, 8.0, 5), synth.ImageSpec(72, 40, "444", 60, 6.0, 32, qtables=3),
             synth.ImageSpec(800, 600, "420", 85, 8.0, 8, qtables=3)]
    encs = [synth.encode(s, truth=True) for s in specs]
    bad = b"\xff\xd8\xff\xc0\x00\x04\x0c\x00\xff\xd9"  # 12-bit SOF: UNSUPPORTED
    b = BatchDecoder(IDCT_MFMA, planes=True)
    errs = b.stage([bad] + [e.data for e in encs])
    assert errs[0] != 0 and not errs[1:].any()
    b.decode()
    errs = b.finish()
    assert errs[0] != 0 and not errs[1:].any()
    for i, (s, e) in enumerate(zip(specs, encs), start=1):
        p = oracle.parse(e.data)
        f64 = oracle.planes(p, e.coefs, 1)
        i32 = oracle.planes(p, e.coefs, 0)
        assert np.array_equal(b.coefficients(i), e.coefs), s
        for c in range(p.num_channels):
            got = b.plane(i, c).astype(int)
            d = np.abs(got - f64[c])
            print("image", i, s, "comp", c, "max", d.max(), "exact", (d == 0).mean(), flush=True)
            assert d.max() <= 1 and (d == 0).mean() > 0.999, (s, c)
            assert np.abs(got - i32[c]).max() <= 1, (s, c)
    b.close()


def _read_pnm(path):
    data = open(path, "rb").read()
    tokens, pos = [], 0
    while len(tokens) < 4:  # magic, width, height, maxval
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        tokens.append(data[pos:end])
        pos = end
    pos += 1  # the single whitespace byte after maxval
    w, h = int(tokens[1]), int(tokens[2])
    ch = 3 if tokens[0] == b"P6" else 1
    assert tokens[0] in (b"P5", b"P6") and int(tokens[3]) == 255
    return np.frombuffer(data[pos:], np.uint8).reshape(h, w, ch)


def test_write_ppm_pgm_equals_oracle(gpu, tmp_path):
    # writeJPG (src/format.cu:218-229): P6 for colour, P5 for gray, of the
    # reader's pixels
    specs = [synth.ImageSpec(96, 64, "420", 80, 6.0, 7), synth.ImageSpec(45, 31, "gray", 70, 6.0, 8),
             synth.ImageSpec(33, 17, "444", 95, 8.0, 2)]
    r = JPGReader(IDCT_INT)
    for k, s in enumerate(specs):
        e = synth.encode(s, truth=True)
        assert r.open(e.data) == 0
        path = str(tmp_path / f"out{k}.{'pgm' if s.sampling == 'gray' else 'ppm'}")
        r.write(path)
        assert np.array_equal(_read_pnm(path), _ref(e)), s
    r.close()


def test_device_frame_views_equal_downloads(gpu):
    # zero-copy torch views of frames and planes in device memory equal the
    # host downloads
    import torch
    specs = [synth.ImageSpec(96, 64, "420", 80, 6.0, 7), synth.ImageSpec(45, 31, "gray", 70, 6.0, 8)]
    encs = [synth.encode(s, truth=True) for s in specs]
    b = BatchDecoder(IDCT_INT, planes=True)
    assert not b.stage([e.data for e in encs]).any()
    b.decode()
    assert not b.finish().any()
    torch.cuda.synchronize()
    for i, (s, e) in enumerate(zip(specs, encs)):
        ptr, n = b.device_ptr(i)
        ch = 1 if s.sampling == "gray" else 3
        assert n == s.width * s.height * ch
        view = device_frame(ptr, s.height, s.width, ch)
        assert view.is_cuda and view.dtype == torch.uint8
        assert np.array_equal(view.cpu().numpy(), b.pixels(i))
        assert np.array_equal(b.pixels(i), _ref(e))
        f = b.info(i)
        pptr, pn = b.device_ptr(i, 2)  # plane 0
        assert pn == f.comp_stride[0] * f.comp_plane_height[0]
        pv = device_frame(pptr, f.comp_plane_height[0], f.comp_stride[0], 1)
        assert np.array_equal(pv.cpu().numpy()[:, :, 0], b.plane(i, 0))
    b.close()
