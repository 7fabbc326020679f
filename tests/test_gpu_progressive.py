"""Progressive JPEG (SOF2) on the GPU, behind CJD_EXT_PROGRESSIVE.

The reference refuses SOF2 (src/format.cu:198-200) with SYNTAX_ERROR; so does
the batch without the flag.  The oracle's progressive decode is pinned in
tests/test_progressive.py (libjpeg baseline/progressive twins, per-scan
point-transform prefixes).  Here k_prog_prepare and k_prog decode libjpeg's
progressive files (the committed fixtures and larger ones written through
Pillow) beside baseline files in the same batch:
* coefficients equal the oracle's;
* integer-iDCT pixels equal the oracle's; FP32 and MFMA planes are within 1
  of the oracle's FP64 planes;
* a second decode of the staged batch, graph replays, the serial entropy
  mode, the decode in parts and the reader API give the same; corrupt
  streams get the oracle's verdict."""
import time

import numpy as np
import pytest
import torch

import oracle
from cudajpeg_amd import ENTROPY_SERIAL, IDCT_FLOAT, IDCT_INT, IDCT_MFMA, BatchDecoder, JPGReader, synth
from cudajpeg_amd._lib import EXT_PROGRESSIVE
from tests.test_progressive import NAMES, pair, pillow_progressive, scan_bounds, scans

pytestmark = pytest.mark.gpu

PR = oracle.EXT_PROGRESSIVE


def _files():
    files = [pair(n)[1] for n in NAMES]
    files += [pillow_progressive(640, 480, 31, q=90),
              pillow_progressive(333, 217, 32, restart_marker_blocks=5, subsampling=1),
              pillow_progressive(501, 77, 33, gray=True, q=75, restart_marker_rows=1),
              pillow_progressive(260, 300, 34, q=100, subsampling=0)]
    base = [synth.encode(synth.ImageSpec(800, 600, "420", 75, 8.0, 35)).data,
            synth.encode(synth.ImageSpec(96, 64, "gray", 50, 6.0, 36)).data]
    is_prog = [True] * len(files) + [False] * len(base)
    return files + base, is_prog


def _check(b, i, data, idct):
    p = oracle.parse(data, PR)
    assert p.error == 0
    q = oracle.decode_scan(p).qcoef
    assert np.array_equal(b.coefficients(i), q), i
    if idct == IDCT_INT:
        assert np.array_equal(b.pixels(i), oracle.colour(p, oracle.planes(p, q, 0))), i
    else:
        f64 = oracle.planes(p, q, 1)
        for c in range(p.num_channels):
            d = np.abs(b.plane(i, c).astype(int) - f64[c])
            assert d.max() <= 1, (i, c)


@pytest.mark.parametrize("idct", [IDCT_INT, IDCT_FLOAT, IDCT_MFMA])
def test_progressive_batch(gpu, idct):
    files, is_prog = _files()
    b = BatchDecoder(idct, tile_bits=512, planes=idct != IDCT_INT)
    errs = b.stage(files)
    assert list(errs) == [1 if x else 0 for x in is_prog]  # the reference's behaviour by default
    b.set_extensions(EXT_PROGRESSIVE)
    assert not b.stage(files).any()
    for rep in range(2):  # the second decode starts again from the staged scans
        b.decode()
        assert not b.finish().any()
        for i, d in enumerate(files):
            print("case progressive", idct, rep, i, flush=True)
            _check(b, i, d, idct)
    b.close()


def test_progressive_large(gpu):
    """A 4096x3072 4:2:0 file without restart markers (one lane per scan)
    and with a restart interval per MCU row (one lane per row of a scan);
    the decode time of each is printed."""
    files = [pillow_progressive(4096, 3072, 41, q=85),
             pillow_progressive(4096, 3072, 41, q=85, restart_marker_rows=1)]
    for k, d in enumerate(files):
        b = BatchDecoder(IDCT_INT)
        b.set_extensions(EXT_PROGRESSIVE)
        assert not b.stage([d]).any()
        b.decode()
        assert not b.finish().any()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.decode()
        assert not b.finish().any()
        torch.cuda.synchronize()
        print("progressive 4096x3072", "restart rows" if k else "no restarts",
              "%.1f ms" % ((time.perf_counter() - t0) * 1e3), flush=True)
        _check(b, 0, d, IDCT_INT)
        b.close()


def test_progressive_serial_entropy(gpu):
    files, _ = _files()
    b = BatchDecoder(IDCT_INT, entropy=ENTROPY_SERIAL)
    b.set_extensions(EXT_PROGRESSIVE)
    assert not b.stage(files).any()
    b.decode()
    assert not b.finish().any()
    for i, d in enumerate(files):
        print("case progressive serial", i, flush=True)
        _check(b, i, d, IDCT_INT)
    b.close()


def test_progressive_graph_replays(gpu):
    files, _ = _files()
    b = BatchDecoder(IDCT_INT, tile_bits=1024)
    b.set_extensions(EXT_PROGRESSIVE)
    assert not b.stage(files).any()
    st = torch.cuda.Stream()
    b.decode(st)
    assert not b.finish(st).any()
    b.capture(st)
    for rep in range(3):
        b.replay(st)
        assert not b.finish(st).any()
        torch.cuda.synchronize()
        for i, d in enumerate(files):
            print("case progressive replay", rep, i, flush=True)
            _check(b, i, d, IDCT_INT)
    b.close()


def test_progressive_reader_api(gpu):
    data = pillow_progressive(333, 217, 32, restart_marker_blocks=5, subsampling=1)
    r = JPGReader(IDCT_INT, keep_coefficients=True)
    assert r.open(data) == 1
    r.close()
    r = JPGReader(IDCT_INT, keep_coefficients=True, extensions=EXT_PROGRESSIVE)
    assert r.open(data) == 0
    p = oracle.parse(data, PR)
    q = oracle.decode_scan(p).qcoef
    assert np.array_equal(r.pixels, oracle.colour(p, oracle.planes(p, q, 0)))
    work = oracle.decode_scan(p, want_work=True).work
    for c in range(p.num_channels):
        assert np.array_equal(r.working_space(c), work[c]), c
    r.close()


def test_progressive_corrupt_streams(gpu):
    """Random bytes of one scan overwritten (markers kept): the batch's
    verdict per image is the oracle's, and decoded images match it."""
    rs = np.random.RandomState(7)
    files = []
    for it in range(120):
        prog = bytearray(pair(NAMES[it % len(NAMES)])[1])
        pp = oracle.parse(bytes(prog), PR)
        sc = scans(pp)
        a, e = scan_bounds(pp, sc[rs.randint(len(sc))])
        for _ in range(rs.randint(1, 4)):
            pos = rs.randint(a, e) if e > a else a
            if pos >= e or prog[pos] == 0xFF or prog[pos - 1] == 0xFF:
                continue
            prog[pos] = min(rs.randint(256), 0xFE)
        files.append(bytes(prog))
    b = BatchDecoder(IDCT_INT)
    b.set_extensions(EXT_PROGRESSIVE)
    assert not b.stage(files).any()
    b.decode()
    errs = b.finish()
    for i, d in enumerate(files):
        p = oracle.parse(d, PR)
        o = oracle.decode_scan(p)
        assert errs[i] == o.error, i
        if not o.error:
            assert np.array_equal(b.coefficients(i), o.qcoef), i
    b.close()


@pytest.mark.parametrize("parts", [1, 2, 3])
def test_progressive_in_decode_parts(gpu, parts):
    """Progressive images between baseline images that give the batch
    hundreds of entropy work items (512-bit tiles, 16 per item), so the
    decode runs in parts: the progressive scans are decoded before the first
    run's entropy launch, whichever run their pixels fall in."""
    base = [synth.encode(synth.ImageSpec(1024, 1024, "420", 90, 8.0, 70 + i)).data for i in range(6)]
    prog = [pair(n)[1] for n in NAMES[:5]] + [pillow_progressive(640, 480, 31, q=90)]
    files = base[:3] + prog[:3] + base[3:] + prog[3:]
    b = BatchDecoder(IDCT_INT, tile_bits=512)
    b.set_tuning(0, 16)
    b.set_overlap(parts)
    b.set_extensions(EXT_PROGRESSIVE)
    assert not b.stage(files).any()
    for rep in range(2):
        b.decode()
        assert not b.finish().any()
        for i, d in enumerate(files):
            print("case progressive parts", parts, rep, i, flush=True)
            _check(b, i, d, IDCT_INT)
    b.close()


def test_progressive_through_serving_loop(gpu):
    """The serving loop with CJD_EXT_PROGRESSIVE in its options decodes
    progressive requests beside baseline ones; without it, submitting a
    progressive file fails with the reference's SYNTAX_ERROR."""
    from cudajpeg_amd.serve import Server
    files = [pair(n)[1] for n in NAMES] + [synth.encode(synth.ImageSpec(96, 64, "420", 80, 6.0, 80)).data]
    srv = Server(max_batch=8, idct=IDCT_INT)
    with pytest.raises(Exception) as exc:
        srv.submit(files[0])
    assert getattr(exc.value, "code", None) == 1
    srv.close()
    srv = Server(max_batch=8, idct=IDCT_INT, extensions=EXT_PROGRESSIVE)
    ids = {srv.submit(d): i for i, d in enumerate(files)}
    got = {}
    deadline = time.time() + 60
    while len(got) < len(files) and time.time() < deadline:
        for c in srv.poll(wait_us=100000):
            i = ids[c.id]
            assert c.error == 0, i
            got[i] = np.asarray(srv.download(c)).copy()
            srv.release(c.id)
    srv.close()
    assert len(got) == len(files)
    for i, d in enumerate(files):
        p = oracle.parse(d, PR)
        ref = oracle.colour(p, oracle.planes(p, oracle.decode_scan(p).qcoef, 0))
        assert np.array_equal(got[i].reshape(-1), ref.reshape(-1)), i
