"""GPU test of the per-kernel device times of a decode (cjd_batch_timings,
cjd_batch_kernel_timings): HIP events around the kernels on the decode
stream, used by the benchmark's roofline."""
import pytest

from cudajpeg_amd import IDCT_INT, BatchDecoder, synth

pytestmark = pytest.mark.gpu


def test_kernel_timings_split_the_entropy_stage(gpu):
    # the per-kernel events lie inside the entropy stage: the tiled decode
    # kernel plus what follows it in the stage never exceed it
    specs = [synth.ImageSpec(320, 240, "420", 80, 6.0, 81), synth.ImageSpec(200, 136, "444", 90, 8.0, 82)]
    b = BatchDecoder(IDCT_INT)
    assert not b.stage([synth.encode(s).data for s in specs]).any()
    b.decode()
    assert not b.finish().any()
    stage, pixel, total = b.timings()
    k, x = b.kernel_timings()
    print("entropy stage", stage, "k_entropy", k, "k_expand", x, "pixel", pixel, "total", total, flush=True)
    assert 0 < k <= stage and 0 <= x <= stage and k + x <= stage + 1e-3
    assert 0 < pixel <= total and stage <= total
    b.close()


def test_graph_capture_needs_a_stream_and_a_graph(gpu):
    # capture on the null stream is refused (PROGRAMMER_ERROR), a replay with
    # nothing captured too; the batch stays usable for direct launches
    from cudajpeg_amd import DecodeError
    b = BatchDecoder(IDCT_INT)
    assert not b.stage([synth.encode(synth.ImageSpec(64, 48, "420", 75, 6.0, 1)).data]).any()
    with pytest.raises(DecodeError) as exc:
        b.capture(None)
    assert exc.value.code == -1
    with pytest.raises(DecodeError):
        b.replay(None)
    b.decode()
    assert not b.finish().any()
    b.close()
