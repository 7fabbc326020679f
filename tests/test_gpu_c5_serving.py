"""Serving parity at the serving workload's shape: requests of the mixed
distribution plus 8192x8192 4:2:0 q85 outliers arrive as a Poisson stream
into the serving loop (continuous batching: flush at 64 requests or 2 ms,
admission by frame memory, several batches in flight).  Each completed frame
must equal the same file decoded alone: bit-exact with the oracle for the
integer iDCT, and bit-identical to a one-image decode by the same kernels for
the MFMA iDCT."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_INT, IDCT_MFMA, BatchDecoder, Server, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def requests():
    rng = np.random.default_rng(5005)
    specs = synth.config_mixed(40, 5000)
    specs += [synth.ImageSpec(8192, 8192, "420", 85, float(rng.uniform(2, 12)), 5500 + k) for k in range(2)]
    specs = [specs[i] for i in rng.permutation(len(specs))]
    return synth.encode_many(specs, truth=True)


def _oracle_rgb(enc):
    p = oracle.parse(enc.data)
    return oracle.colour(p, oracle.planes(p, enc.coefs, 0))


def _alone(enc, idct):
    b = BatchDecoder(idct)
    assert b.stage([enc.data])[0] == 0
    b.decode()
    assert b.finish()[0] == 0
    out = b.pixels(0)
    b.close()
    return out


def _serve(encs, idct, rate):
    """Submit every request at Poisson times (rate per second), collect every
    completion with its host copy of the frame; returns {request: (c, rgb)}."""
    rng = np.random.default_rng(11)
    gaps = rng.exponential(1.0 / rate, len(encs))
    srv = Server(max_batch=64, max_wait_us=2000, idct=idct)
    ids = {}
    got = {}

    def take(wait_us):
        for c in srv.poll(256, wait_us):
            k = ids[c.id]
            got[k] = (c, srv.download(c) if c.error == 0 else None)
            srv.release(c.id)

    t = time.perf_counter()
    for k, e in enumerate(encs):
        t += gaps[k]
        while time.perf_counter() < t:
            take(0)
        ids[srv.submit(e.data)] = k
    for _ in range(4000):
        if len(got) == len(encs):
            break
        take(5000)
    srv.close()
    assert len(got) == len(encs)
    return got


def test_c5_serving_int_equals_oracle_alone(gpu, requests):
    print("case C5 integer iDCT:", len(requests), "requests", flush=True)
    got = _serve(requests, IDCT_INT, 2000.0)
    sizes = sorted({c.batch_size for c, _ in got.values()})
    print("batch sizes seen", sizes, flush=True)
    with ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(_oracle_rgb, requests))
    for k, e in enumerate(requests):
        c, rgb = got[k]
        assert c.error == 0, (k, e.spec, c.error)
        assert (c.width, c.height) == (e.spec.width, e.spec.height)
        assert c.t_arrival_ns <= c.t_launch_ns <= c.t_ready_ns
        assert np.array_equal(rgb, refs[k]), (k, e.spec)
    # requests were batched together (continuous batching), outliers included
    assert max(sizes) > 1


def test_c5_serving_mfma_equals_alone(gpu, requests):
    print("case C5 MFMA iDCT:", len(requests), "requests", flush=True)
    got = _serve(requests, IDCT_MFMA, 2000.0)
    for k, e in enumerate(requests):
        c, rgb = got[k]
        assert c.error == 0, (k, e.spec, c.error)
        assert np.array_equal(rgb, _alone(e, IDCT_MFMA)), (k, e.spec)
