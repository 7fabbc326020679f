"""GPU tests of the serving loop (cjd_server_*): frames returned through the
loop equal the oracle's pixels, admission by frame memory, and a replayed
Poisson arrival trace."""
import numpy as np
import pytest

import oracle
from cudajpeg_amd import IDCT_INT, Server, synth

pytestmark = pytest.mark.gpu

SPECS = [synth.ImageSpec(64 + 8 * k, 48 + 6 * k, ["420", "444", "gray", "422"][k % 4], 60 + 3 * k, 6.0, 50 + k)
         for k in range(12)]


@pytest.fixture(scope="module")
def encoded():
    return [synth.encode(s, truth=True) for s in SPECS]


def _ref(enc):
    p = oracle.parse(enc.data)
    return oracle.colour(p, oracle.planes(p, enc.coefs, 0))


def _drain(srv, n, timeout_polls=2000):
    got = []
    for _ in range(timeout_polls):
        got += srv.poll(256, 5000)
        if len(got) >= n:
            break
    return got


def test_server_frames_equal_oracle(gpu, encoded):
    srv = Server(max_batch=5, max_wait_us=500, idct=IDCT_INT)
    ids = {srv.submit(e.data): k for k, e in enumerate(encoded)}
    got = _drain(srv, len(encoded))
    assert len(got) == len(encoded)
    for c in got:
        k = ids[c.id]
        print("request", k, SPECS[k], "batch", c.batch_size, flush=True)
        assert c.error == 0 and c.frame
        assert (c.width, c.height) == (SPECS[k].width, SPECS[k].height)
        assert c.t_arrival_ns <= c.t_launch_ns <= c.t_ready_ns
        assert np.array_equal(srv.download(c), _ref(encoded[k])), SPECS[k]
        srv.release(c.id)
    srv.close()


def test_server_admission_by_frame_memory(gpu, encoded):
    # a budget of two of the largest frames: every request still completes,
    # in batches whose frames fit the budget (frames are released as they
    # are checked, which lets the loop admit more); a request larger than
    # the budget fails with the reference's out-of-memory code
    big = max(s.width * s.height * 3 for s in SPECS)
    budget = 2 * big
    srv = Server(max_batch=64, max_wait_us=200, frame_budget=budget, idct=IDCT_INT)
    ids = {srv.submit(e.data): k for k, e in enumerate(encoded)}
    huge = synth.encode(synth.ImageSpec(400, 400, "444", 80, 6.0, 99)).data
    hid = srv.submit(huge)
    per_launch = {}
    seen = 0
    for _ in range(2000):
        for c in srv.poll(256, 5000):
            seen += 1
            if c.id == hid:
                assert c.error == 3
            else:
                assert c.error == 0
                assert np.array_equal(srv.download(c), _ref(encoded[ids[c.id]]))
                per_launch[c.t_launch_ns] = per_launch.get(c.t_launch_ns, 0) + c.frame_bytes
            srv.release(c.id)
        if seen == len(encoded) + 1:
            break
    srv.close()
    assert seen == len(encoded) + 1
    assert len(per_launch) > 1 and max(per_launch.values()) <= budget


def test_server_rejects_bad_input(gpu):
    srv = Server()
    with pytest.raises(Exception):
        srv.submit(b"\xff\xd8\xff\xd9")
    srv.close()


def test_server_replay_poisson(gpu, encoded):
    rng = np.random.default_rng(7)
    n = 400
    files = [encoded[k % len(encoded)].data for k in range(n)]
    arrival = np.cumsum(rng.exponential(1e9 / 4000.0, n)).astype(np.int64)  # 4000 requests/s
    srv = Server(max_batch=64, max_wait_us=2000)
    lat, err, wall = srv.replay(files, arrival)
    srv.close()
    print("replay p50/p99 ms", np.percentile(lat, 50) / 1e6, np.percentile(lat, 99) / 1e6, "wall s", wall / 1e9)
    assert not err.any()
    assert (lat > 0).all() and wall > 0
