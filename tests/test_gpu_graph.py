"""A batch's decode captured as a HIP graph and replayed (cjd_batch_capture /
cjd_batch_replay).

Every replay is checked against the direct decode of the same batch, with the
program's own work between replays: a direct decode of another batch on
another stream, torch kernels, synchronisations, and direct decodes of the
same batch.  Covered: the decode in 2 parts (second stream joined into the
graph by events) and in one part, the MFMA and integer forms, the redo path
after a replay (a wait limit of 0 makes images unresolved), re-staging files
of the same shape (the graph stays valid) and of another shape (the replay is
refused with UNSUPPORTED_ERROR)."""
import numpy as np
import pytest
import torch

from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, synth
from cudajpeg_amd.decoder import DecodeError

pytestmark = pytest.mark.gpu


def _outputs(d, n):
    return [d.pixels(i) for i in range(n)], [d.coefficients(i) for i in range(n)]


def _same(a, b, what):
    for i, (x, y) in enumerate(zip(a[0], b[0])):
        assert np.array_equal(x, y), (what, "pixels", i)
    for i, (x, y) in enumerate(zip(a[1], b[1])):
        assert np.array_equal(x, y), (what, "coefficients", i)


@pytest.mark.parametrize("idct,parts", [(IDCT_MFMA, 2), (IDCT_INT, 1), (IDCT_INT, 2)])
def test_graph_replays_equal_direct_decodes(gpu, idct, parts):
    specs = synth.config_mixed(64, 8500 + parts)
    encs = synth.encode_many(specs, truth=True)
    files = [e.data for e in encs]
    other = synth.encode_many(synth.config_mixed(16, 8600))
    d = BatchDecoder(idct, tile_bits=8192)
    d.set_overlap(parts)
    d.set_tuning(0, 128)
    assert not d.stage(files).any()
    st, st2 = torch.cuda.Stream(), torch.cuda.Stream()
    d.decode(st)
    assert not d.finish(st).any()
    ref = _outputs(d, len(files))
    for i, e in enumerate(encs):
        assert np.array_equal(ref[1][i], e.coefs), i
    o = BatchDecoder(IDCT_INT)
    assert not o.stage([e.data for e in other]).any()
    d.capture(st)
    x = torch.randn(1 << 22, device="cuda")
    for rep in range(5):
        print("case graph", idct, parts, "replay", rep, flush=True)
        o.decode(st2)                      # another batch's direct decode meanwhile
        d.replay(st)
        with torch.cuda.stream(st2):
            y = (x * 2.0).sum()            # torch work of the program's own
        assert not d.finish(st).any()
        assert not o.finish(st2).any()
        torch.cuda.synchronize()
        assert float(y) == pytest.approx(float(x.sum()) * 2.0, rel=1e-4)
        _same(_outputs(d, len(files)), ref, ("replay", rep))
        assert d.protocol_redos() == 0
        if rep == 2:                       # a direct decode of the batch between replays
            d.decode(st)
            assert not d.finish(st).any()
            _same(_outputs(d, len(files)), ref, "direct between replays")
    # the same files staged again: the graph still matches the batch
    assert not d.stage(files).any()
    d.replay(st)
    assert not d.finish(st).any()
    _same(_outputs(d, len(files)), ref, "replay after re-staging")
    # other files: the replay is refused, the batch decodes directly
    assert not d.stage(files[:5]).any()
    with pytest.raises(DecodeError) as ei:
        d.replay(st)
    assert ei.value.code == 2
    d.decode(st)
    assert not d.finish(st).any()
    for i in range(5):
        assert np.array_equal(d.coefficients(i), encs[i].coefs), i
    d.close()
    o.close()


def test_graph_replay_redoes_unresolved_images(gpu):
    # a wait limit of 0 (forced from the host) leaves images unresolved in
    # every replay; the host's redo after each replay makes them exact
    specs = [synth.ImageSpec(1200, 900, "420", 63, 8.9, 5), synth.ImageSpec(800, 600, "gray", 90, 12.0, 3),
             synth.ImageSpec(1000, 700, "422", 80, 6.0, 9, restart_interval=7)]
    encs = [synth.encode(s, truth=True) for s in specs]
    d = BatchDecoder(IDCT_INT, tile_bits=2048)
    d.set_tuning(0, 4)
    d.set_wait_limit(0)
    assert not d.stage([e.data for e in encs]).any()
    st = torch.cuda.Stream()
    d.capture(st)
    for rep in range(3):
        print("case graph redo replay", rep, flush=True)
        d.replay(st)
        assert not d.finish(st).any()
        for i, e in enumerate(encs):
            assert np.array_equal(d.coefficients(i), e.coefs), (rep, i)
        assert d.protocol_redos() == 0
    assert d.fallbacks() > 0
    d.close()
