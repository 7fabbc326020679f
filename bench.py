"""Benchmark of the MI355X baseline-JPEG decoder.

Workload (one "step"): decode the 256-image mixed batch -- areas log-uniform
in 256x256..4096x3072, aspect 1:1/4:3/16:9, 70% 4:2:0 / 20% 4:4:4 / 10% gray,
quality U{60..95}, noise U[2,12] -- with the scans already resident in HBM
and the RGB/gray frames left in HBM.

Reported (all by default; --no-plain, --no-serve, --no-progressive, --no-cpu
skip sections):
  value              batched form, Mpixel/s (W*H): each decode runs in
                     --parts runs of images (2): each run's pixel kernel on a
                     second stream beside the next runs' entropy decodes,
                     the last run's after its own; --in-flight
                     N > 1 lets N decoders of the batch take the steps in turn
                     on their own streams (off by default)
  kernel_ms          device time per kernel (HIP events on the decode stream,
                     the kernels run one after the other)
  roofline           dominant kernel, achieved algorithmic GB/s vs 8 TB/s;
                     traffic = its HBM bytes per launch from the committed
                     FETCH_SIZE/WRITE_SIZE summary of this workload and these
                     kernels, null when there is none
  plain_mpix_s       plain form: one image per launch (entropy + pixel
                     kernels), FP32 VALU iDCT, per-size entropy parameters
                     from a run-time autotune; ratio = batched / plain;
                     p50/p99_ms = per-image decode latency of the plain form
                     (enqueue to stream synchronised)
  serving            the serving loop on the serving mix (mixed batch + 5%
                     8192x8192 4:2:0 q85): peak rate, then Poisson arrivals at
                     each --serve-rates rate below peak and at 70% of peak:
                     p50/p99/p99.9
                     latency (submit to frame ready in device memory) and
                     sustained Mpixel/s
  cpu_baseline       the reference CPUonly decoder (built from its sources
                     into oracle/_ref) on a bounded sample, one host core
                     (a worker process pinned to one CPU); all_cores: the
                     whole batch over --cpu-workers (16) processes started
                     together; the host CPU model is stated

  progressive        progressive files (CJD_EXT_PROGRESSIVE, written by
                     libjpeg through Pillow; progressive_bench): one
                     4096x3072 file without and with restart intervals, 64
                     restart-free 1024x768 files, beside libjpeg's one-core
                     decode of the same files

--progressive runs only that section and prints one JSON line of its own.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def check(errs, what: str) -> None:
    """Stop with the failing images and their error codes when a decode
    reported errors."""
    errs = np.asarray(errs)
    bad = np.flatnonzero(errs)
    if bad.size:
        raise SystemExit(f"bench.py: {what}: {bad.size} image(s) with errors: "
                         + ", ".join(f"#{i} code {int(errs[i])}" for i in bad[:16]))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_PINNED_WORKER = r"""
import json, os, sys, time
sys.path.insert(0, sys.argv[1])
core, budget, kind, paths = int(sys.argv[2]), float(sys.argv[3]), sys.argv[4], sys.argv[5:]
os.sched_setaffinity(0, {core})
import oracle
times = []
for p in paths:
    if kind == "reference":
        t = float(oracle.ref_cpuonly_time([p], 1)[0, 0])
    else:
        data = open(p, "rb").read()
        t0 = time.perf_counter()
        err, _ = oracle.decode(data, 0)
        t = time.perf_counter() - t0
        assert err == 0
    times.append(t)
    if sum(times) >= budget:
        break
print(json.dumps(times))
"""


def cpu_baseline(encs, budget_s: float):
    """Reference CPUonly (oracle/_ref) or, if it was not built, the oracle's
    C restatement, on a bounded sample of the same workload, in a worker
    process pinned to one CPU (taskset -c <cpu>)."""
    import subprocess
    import oracle
    kind = "reference" if os.path.exists(os.path.join(oracle.REF_DIR, "libcpuonly.so")) else "port"
    tmp = tempfile.mkdtemp(prefix="cjd_bench_")
    order = sorted(range(len(encs)), key=lambda i: encs[i].spec.pixels)
    # a spread of sizes: every 8th image by area, then the others (each once)
    sample = [i for i in order[::8] + [order[k] for k in range(len(order)) if k % 8]
              if not (kind == "reference" and encs[i].spec.restart_interval)]
    assert len(set(sample)) == len(sample)
    paths = []
    for i in sample:
        path = os.path.join(tmp, f"{i}.jpg")
        with open(path, "wb") as f:
            f.write(encs[i].data)
        paths.append(path)
    core = min(os.sched_getaffinity(0))
    out = subprocess.run([sys.executable, "-c", _PINNED_WORKER, ROOT, str(core), str(budget_s), kind] + paths,
                         capture_output=True, text=True, timeout=600)
    for path in paths:
        os.unlink(path)
    os.rmdir(tmp)
    if out.returncode:
        raise SystemExit(f"bench.py: CPU baseline worker failed: {out.stderr[-500:]}")
    times = json.loads(out.stdout.strip().splitlines()[-1])
    n = len(times)
    px = sum(encs[i].spec.pixels for i in sample[:n])
    done_t = sum(times)
    return {"value": round(px / done_t / 1e6, 2), "unit": "Mpix/s", "cores": 1, "kind": kind,
            "cpu_model": cpu_model(), "pinned_cpu": core,
            "p50_ms": round(float(np.percentile(times, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(times, 99)) * 1e3, 3),
            "sample": f"{n} of the 256 mixed images (spread of sizes), decoded once each by the "
                      f"{'reference CPUonly openJPG' if kind == 'reference' else 'oracle C restatement'}, "
                      f"file in page cache, one process pinned to CPU {core}, {done_t:.1f} s total, "
                      f"{px / 1e6:.0f} Mpix"}


_CPU_WORKER = r"""
import os, sys, time
sys.path.insert(0, sys.argv[1])
import oracle
go, out, reps, paths = sys.argv[2], sys.argv[3], int(sys.argv[4]), sys.argv[5:]
open(out + ".ready", "w").close()
t_end = time.time() + 120
while not os.path.exists(go):
    if time.time() > t_end:
        sys.exit(3)
    time.sleep(0.001)
t0 = time.perf_counter()
oracle.ref_cpuonly_time(paths, reps)
with open(out, "w") as f:
    f.write(repr(time.perf_counter() - t0))
"""


def cpu_all_cores(encs, workers: int, reps: int):
    """The reference CPUonly on every image of the batch (restart-interval
    files aside), `reps` times each, split over `workers` independent
    processes (greedy by pixels) that start together; Mpix/s over the
    slowest process's wall time."""
    import subprocess
    import oracle
    if not os.path.exists(os.path.join(oracle.REF_DIR, "libcpuonly.so")):
        return None
    tmp = tempfile.mkdtemp(prefix="cjd_bench_mc_")
    parts = [[] for _ in range(workers)]
    load = [0] * workers
    total_px = 0
    for i in sorted(range(len(encs)), key=lambda i: -encs[i].spec.pixels):
        e = encs[i]
        if e.spec.restart_interval:
            continue
        path = os.path.join(tmp, f"{i}.jpg")
        with open(path, "wb") as f:
            f.write(e.data)
        k = load.index(min(load))
        parts[k].append(path)
        load[k] += e.spec.pixels
        total_px += e.spec.pixels * reps
    go = os.path.join(tmp, "go")
    root = os.path.dirname(os.path.abspath(__file__))
    procs = []
    for k, ps in enumerate(parts):
        if ps:
            out = os.path.join(tmp, f"w{k}")
            procs.append((out, subprocess.Popen([sys.executable, "-c", _CPU_WORKER, root, go, out, str(reps)] + ps)))
    t_end = time.time() + 120
    while not all(os.path.exists(o + ".ready") for o, _ in procs):
        if time.time() > t_end or any(p.poll() not in (None, 0) for _, p in procs):
            for _, p in procs:
                p.kill()
            raise SystemExit("bench.py: CPU workers did not start")
        time.sleep(0.01)
    open(go, "w").close()
    for _, p in procs:
        if p.wait(timeout=300) != 0:
            raise SystemExit("bench.py: a CPU worker failed")
    wall = max(float(open(o).read()) for o, _ in procs)
    for name in os.listdir(tmp):
        os.unlink(os.path.join(tmp, name))
    os.rmdir(tmp)
    return {"value": round(total_px / wall / 1e6, 2), "unit": "Mpix/s", "cores": len(procs),
            "sample": f"the batch's {sum(len(ps) for ps in parts)} images without restart intervals, "
                      f"{reps} decodes each, over {len(procs)} processes of the reference CPUonly "
                      f"started together, {wall:.2f} s wall"}


def serve_bench(args):
    """Serving loop on the serving mix: the mixed-batch distribution plus 5%
    8192x8192 4:2:0 q85 images (a pool of 252 files sampled with
    replacement).  Peak = a burst of requests decoded flat out; then Poisson
    arrivals at each of --serve-rates and at 70% of peak, each for about
    --serve-seconds (at most 6000 requests); latency from submit to the frame
    ready in device memory, sustained Mpixel/s = pixels completed / wall.
    Every phase logs its progress (the loop's counters every 2 s); when no
    request completes for --serve-stall seconds while some are outstanding,
    the counters are logged and the process ends (exit status 4)."""
    from cudajpeg_amd import Server, synth
    rng = np.random.default_rng(args.seed + 1)
    t0 = time.time()
    specs = synth.config_mixed(240, args.seed + 1000)
    specs += [synth.ImageSpec(8192, 8192, "420", 85, float(rng.uniform(2, 12)), args.seed + 2000 + k)
              for k in range(12)]
    order = rng.permutation(len(specs))
    specs = [specs[i] for i in order]
    files = [e.data for e in synth.encode_many(specs)]
    px = np.array([s.pixels for s in specs], np.float64)
    log(f"serving pool: {len(files)} files, {sum(map(len, files)) / 1e6:.0f} MB, encoded in {time.time() - t0:.1f} s")
    srv = Server(max_batch=64, max_wait_us=2000)
    log("serving loop created")

    def watch(what):
        last = {"completed": -1, "t": time.time()}

        def progress(st):
            log(f"{what}: {st}")
            now = time.time()
            if st["completed"] != last["completed"]:
                last["completed"], last["t"] = st["completed"], now
            elif st["submitted"] > st["completed"] and now - last["t"] > args.serve_stall:
                log(f"{what}: no request completed for {now - last['t']:.0f} s with "
                    f"{st['submitted'] - st['completed']} outstanding; ending the serving run")
                os._exit(4)
        return progress

    # peak: a burst of 8 x the pool (long enough that the first batch's
    # staging and the last one's tail are a small part of it), after the same
    # burst once untimed, in which every slot grows its buffers to the mix
    n_peak = 8 * len(files)
    burst = [files[k % len(files)] for k in range(n_peak)]
    lat, err, wall = srv.replay(burst, np.zeros(n_peak, np.int64), progress=watch("serving warm-up burst"))
    check(err, "serving warm-up burst")
    log(f"serving warm-up burst: {n_peak} requests in {wall / 1e9:.3f} s")
    lat, err, wall = srv.replay(burst, np.zeros(n_peak, np.int64), progress=watch("serving burst"))
    check(err, "serving burst")
    peak_req_s = n_peak / (wall / 1e9)
    peak_mpix_s = float(sum(px[k % len(files)] for k in range(n_peak)) / (wall / 1e9) / 1e6)
    log(f"serving peak: {n_peak} requests in {wall / 1e9:.3f} s, {peak_req_s:.0f} req/s, {peak_mpix_s:.0f} Mpix/s")

    def run(rate, seconds):
        n = int(min(6000, max(64, rate * seconds)))
        log(f"serving at {rate:.0f} req/s: {n} requests")
        arrival = np.cumsum(rng.exponential(1e9 / rate, n)).astype(np.int64)
        pick = rng.integers(0, len(files), n)
        lat, err, wall = srv.replay([files[int(k)] for k in pick], arrival,
                                    progress=watch(f"serving at {rate:.0f} req/s"))
        check(err, f"serving at {rate:.0f} req/s")
        log(f"serving at {rate:.0f} req/s, loop counters so far: {srv.stats()}")
        ms = lat / 1e6
        return {"offered_req_s": round(rate, 1), "above_peak": bool(rate >= peak_req_s), "requests": n,
                "seconds": round(wall / 1e9, 3),
                "sustained_mpix_s": round(float(px[pick].sum()) / (wall / 1e9) / 1e6, 1),
                "p50_ms": round(float(np.percentile(ms, 50)), 3), "p99_ms": round(float(np.percentile(ms, 99)), 3),
                "p999_ms": round(float(np.percentile(ms, 99.9)), 3)}

    # rates below peak first, then 70% of peak, then the rates at or above
    # peak: those only grow the queue for as long as the arrivals last (at most
    # 6000 requests), so their latencies measure queueing, and their sustained
    # rate is the peak
    sweep = []
    for rate in sorted(r for r in args.serve_rates if r < peak_req_s):
        sweep.append(run(rate, args.serve_seconds))
        log(f"serving {rate:.0f} req/s: {sweep[-1]}")
    at70 = run(0.7 * peak_req_s, max(args.serve_seconds, 3.0))
    log(f"serving at 70% of peak: {at70}")
    for rate in sorted(r for r in args.serve_rates if r >= peak_req_s):
        sweep.append(run(rate, args.serve_seconds))
        log(f"serving {rate:.0f} req/s (at or above peak): {sweep[-1]}")
    counters = srv.stats()
    srv.close()
    return {"request_mix": "240 mixed-batch images + 12 8192x8192 4:2:0 q85 (5%), sampled with replacement",
            "peak_req_s": round(peak_req_s, 1), "peak_mpix_s": round(peak_mpix_s, 1),
            "at_70pct_of_peak": at70, "sweep": sweep, "max_batch": 64, "max_wait_us": 2000,
            "redone_requests": counters["redone"], "batches": counters["batches"],
            "latency": "submit to frame ready in device memory (host monotonic clock)"}


def serve_child(args):
    """The serving section in a child process of its own (bench.py
    --serve-only): its progress lines go to this process's stderr, and it is
    ended after --serve-timeout seconds, so that a serving run that does not
    finish cannot keep the rest of the bench from reporting."""
    import subprocess
    fd, out = tempfile.mkstemp(prefix="cjd_serve_", suffix=".json")
    os.close(fd)
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-only", out, "--seed", str(args.seed),
           "--serve-rates", ",".join(str(r) for r in args.serve_rates), "--serve-seconds", str(args.serve_seconds),
           "--serve-stall", str(args.serve_stall)]
    p = subprocess.Popen(cmd, stdout=sys.stderr)
    try:
        rc = p.wait(timeout=args.serve_timeout)
    except subprocess.TimeoutExpired:
        p.kill()
        p.wait()
        rc = None
    try:
        res = json.load(open(out)) if rc == 0 else None
    except (OSError, ValueError):
        res = None
    os.unlink(out)
    if res is None:
        why = (f"did not finish within {args.serve_timeout:.0f} s" if rc is None
               else "stalled (no request completed)" if rc == 4 else f"exited with status {rc}")
        log(f"serving run {why}")
        return {"error": f"the serving run {why}; its progress lines are in the bench log"}
    return res


def progressive_bench(args) -> dict:
    """Progressive files (CJD_EXT_PROGRESSIVE), written by libjpeg through
    Pillow with its standard ten-scan script: a 4096x3072 4:2:0 q85 file
    without restart markers and with a restart interval per MCU row, and 64
    restart-free 1024x768 files in one batch.  Decode time (stage excluded,
    decode + finish, best of 3) beside libjpeg's own decode of the same
    files through Pillow on one host core."""
    import io
    import importlib.util

    import torch
    from PIL import Image

    from cudajpeg_amd import IDCT_INT, BatchDecoder
    from cudajpeg_amd._lib import EXT_PROGRESSIVE

    spec = importlib.util.spec_from_file_location("make_progressive",
                                                  os.path.join(ROOT, "tests", "golden", "make_progressive.py"))
    mp = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mp)

    def encode(w, h, seed, **opts):
        b = io.BytesIO()
        Image.fromarray(mp.pixels(w, h, seed)).save(b, "JPEG", quality=85, progressive=True, **opts)
        return b.getvalue()

    cases = [("4096x3072, no restarts", [encode(4096, 3072, 41)], 4096 * 3072),
             ("4096x3072, restart per MCU row", [encode(4096, 3072, 41, restart_marker_rows=1)], 4096 * 3072),
             ("64 x 1024x768, no restarts", [encode(1024, 768, 100 + i) for i in range(64)], 64 * 1024 * 768)]
    out = {}
    for name, files, px in cases:
        b = BatchDecoder(IDCT_INT)
        b.set_extensions(EXT_PROGRESSIVE)
        check(b.stage(files), name)
        b.decode()
        check(b.finish(), name)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.decode()
            check(b.finish(), name)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        b.close()
        t0 = time.perf_counter()
        for f in files:
            Image.open(io.BytesIO(f)).load()
        cpu = time.perf_counter() - t0
        out[name] = {"gpu_ms": round(min(ts) * 1e3, 2), "gpu_mpix_s": round(px / min(ts) / 1e6, 1),
                     "libjpeg_1core_ms": round(cpu * 1e3, 1), "libjpeg_1core_mpix_s": round(px / cpu / 1e6, 1)}
        log(f"progressive {name}: {out[name]}")
    return out


def measured_traffic(kernel: str, args) -> int | None:
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE/WRITE_SIZE
    summary of this same workload (profiles/hbm_traffic_mixed256.md, counters
    scaled by a copy kernel of known size); None for any other workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_mixed256.md")
    if args.images != 256 or args.seed != 4000 or not os.path.exists(path):
        return None
    text = open(path).read()
    doc = json.loads(text[text.rfind("```json") + 7:text.rfind("```")])
    if doc.get("entropy_params") != args.entropy_params:
        return None  # the summary is of another tiling
    return doc["traffic_bytes"].get(kernel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--images", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4000)
    # the batched form's entropy parameters: by default the per-bucket table
    # the batched autotuner measured (cudajpeg_amd/tune/mi355x_batched.json);
    # --autotune measures one now; --tile-bits / --item-tiles set one tiling
    # for every image instead
    ap.add_argument("--tile-bits", type=int, default=0, help="one tile length for every image (0 = the table)")
    ap.add_argument("--item-tiles", type=int, default=0, help="tiles per entropy work item for every image "
                                                               "(0 = the table)")
    ap.add_argument("--autotune", action="store_true",
                    help="measure the per-bucket table of the batched form now (autotune_batched) and use it")
    ap.add_argument("--ck-bits", type=int, default=0, help="checkpoint spacing of the tiled decoder (0 = default)")
    ap.add_argument("--batched-idct", default="mfma", choices=["int", "float", "mfma"])
    ap.add_argument("--parts", type=int, default=2,
                    help="decode in parts: each run's pixel kernel beside the next runs' entropy (1 = off)")
    ap.add_argument("--in-flight", type=int, default=1,
                    help="decoders of the batch taking the timed steps in turn, one stream each")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="processes for the all-cores CPU baseline (the host's share of cores; 1 = off)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain form (autotune, then one image per launch) and its latencies")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop load test")
    ap.add_argument("--no-progressive", action="store_true", help="skip the progressive-file timings")
    ap.add_argument("--no-graph", action="store_true", help="skip the timing of graph replays of the decode")
    ap.add_argument("--serve-rates", type=lambda v: [float(x) for x in v.split(",")],
                    default=[500.0, 1000.0, 2000.0, 4000.0, 8000.0], help="Poisson rates of the sweep, req/s")
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="seconds of arrivals per swept rate")
    ap.add_argument("--serve-timeout", type=float, default=240.0,
                    help="seconds after which the serving run (a child process) is ended")
    ap.add_argument("--serve-stall", type=float, default=45.0,
                    help="seconds without a completed request (some outstanding) that end the serving run")
    ap.add_argument("--serve-only", metavar="PATH", default=None,
                    help="only run the serving section and write its result (JSON) to PATH")
    ap.add_argument("--progressive", action="store_true",
                    help="only time progressive files (CJD_EXT_PROGRESSIVE; needs Pillow) and print them as JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain read+write elementwise kernel over MIB MiB (reads MIB, writes MIB) "
                         "three times: the reference point for scaling FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every image decodes on one device, and the multi-GPU "
              "design (image shards / bit-range split) is documented, not run. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from cudajpeg_amd import IDCT_FLOAT, IDCT_INT, IDCT_MFMA, BatchDecoder, synth
    from cudajpeg_amd._lib import UNSTUFF_DEVICE, UNSTUFF_HOST

    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    if args.serve_only:
        res = serve_bench(args)
        with open(args.serve_only, "w") as f:
            json.dump(res, f)
        return
    if args.progressive:
        print(json.dumps({"progressive": progressive_bench(args)}), flush=True)
        return
    if args.copy_calibration:
        n = args.copy_calibration << 18  # int32 elements
        x = torch.randint(0, 1 << 30, (n,), dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.bitwise_not(x, out=y)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB read + {args.copy_calibration} MiB written per launch")
        return
    dev = torch.cuda.current_device()
    stream = torch.cuda.Stream()

    t0 = time.time()
    specs = synth.config_mixed(args.images, args.seed)
    encs = synth.encode_many(specs)
    files = [e.data for e in encs]
    total_px = sum(s.pixels for s in specs)
    log(f"encoded {len(files)} images, {total_px / 1e6:.0f} Mpix, {sum(map(len, files)) / 1e6:.0f} MB "
        f"in {time.time() - t0:.1f}s")

    idct = {"int": IDCT_INT, "float": IDCT_FLOAT, "mfma": IDCT_MFMA}[args.batched_idct]
    # ---------------------------------------------------------------- batched
    from cudajpeg_amd.autotune import BATCHED_TABLE_MI355X, autotune_batched, load_table
    table = []
    if args.tile_bits or args.item_tiles:
        args.entropy_params = f"uniform {args.tile_bits or 4096} x {args.item_tiles or 256}"
    elif args.autotune:
        t_at = time.time()
        table = autotune_batched(files, idct, args.parts, reps=5)
        args.entropy_params = "autotuned now: " + json.dumps([list(t) for t in table])
        log(f"batched autotune ({time.time() - t_at:.1f} s): {table}")
    else:
        table = load_table(BATCHED_TABLE_MI355X)
        args.entropy_params = "table " + os.path.relpath(BATCHED_TABLE_MI355X, ROOT)

    def batched_decoder():
        d = BatchDecoder(idct, tile_bits=args.tile_bits)
        d.set_overlap(args.parts)
        if args.item_tiles or args.ck_bits:
            d.set_tuning(args.ck_bits, args.item_tiles)
        if table:
            d.set_tune_table(table)
        return d

    b = batched_decoder()
    check(b.stage(files), "stage")
    # host side of the boundary: parse + unstuff + plan + upload of the scans
    def time_stage():
        ms = []
        for _ in range(3):
            t1 = time.perf_counter()
            b.stage(files)
            ms.append((time.perf_counter() - t1) * 1e3)
        return float(np.median(ms))

    # the scans unstuffed by the host planner, then (the default) on the device
    b.set_unstuff(UNSTUFF_HOST)
    stage_ms_host = time_stage()
    b.set_unstuff(UNSTUFF_DEVICE)
    stage_ms = time_stage()
    log(f"staging: {stage_ms:.2f} ms with device unstuffing, {stage_ms_host:.2f} ms with host unstuffing")
    tot = b.totals()
    with torch.cuda.stream(stream):
        b.decode(stream)
        check(b.finish(stream), "first decode")
        ms_ent, ms_pix, ms_all = b.timings()
        log(f"batched first decode: entropy {ms_ent:.2f} ms, pixel {ms_pix:.2f} ms, fallbacks {b.fallbacks()}")
        # one step = one decode of the batch (its launches, enqueued directly);
        # first with one batch in flight, then with --in-flight decoders of the
        # same batch taking the steps in turn, each on its own stream, so that
        # one batch's entropy kernel runs beside the previous one's pixel kernel
        for _ in range(args.warmup):
            b.decode(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            b.decode(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        t_single = (time.perf_counter() - t_start) / args.steps
        decs, streams = [b], [stream]
        for _ in range(args.in_flight - 1):
            d = batched_decoder()
            check(d.stage(files), f"stage of decoder {len(decs)}")
            decs.append(d)
            streams.append(torch.cuda.Stream())
        n_fl = len(decs)
        for k in range(args.warmup * n_fl):
            decs[k % n_fl].decode(streams[k % n_fl])
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for k in range(args.steps):
            decs[k % n_fl].decode(streams[k % n_fl])
        for st in streams:
            st.synchronize()
        torch.cuda.synchronize()
        t_batched = (time.perf_counter() - t_start) / args.steps
        for k, (d, st) in enumerate(zip(decs[1:], streams[1:])):
            check(d.finish(st), f"timed steps, decoder {k + 1} (fallbacks {d.fallbacks()})")
            d.close()
        # per-kernel device times (events on the decode stream), averaged,
        # with the two kernels one after the other so that each time is the
        # kernel's own
        b.set_overlap(1)
        ent, pix, kent, kexp = [], [], [], []
        for _ in range(args.steps):
            b.decode(stream)
            a, c, _ = b.timings()
            ka, kx = b.kernel_timings()
            ent.append(a)
            pix.append(c)
            kent.append(ka)
            kexp.append(kx)
        check(b.finish(stream), f"timed steps, decoder 0 (fallbacks {b.fallbacks()})")
        b.set_overlap(args.parts)
        # the same steps as replays of the decode captured as one HIP graph
        t_graph = None
        if not args.no_graph:
            b.capture(stream)
            for _ in range(args.warmup):
                b.replay(stream)
            check(b.finish(stream), "graph warm-up")
            torch.cuda.synchronize()
            t_start = time.perf_counter()
            for _ in range(args.steps):
                b.replay(stream)
            stream.synchronize()
            torch.cuda.synchronize()
            t_graph = (time.perf_counter() - t_start) / args.steps
            check(b.finish(stream), "graph replays")
            log(f"graph replays: {t_graph * 1e3:.2f} ms/step ({total_px / t_graph / 1e9:.1f} Gpix/s)")
    batched_mpix = total_px / t_batched / 1e6
    # staging pipelined with decoding: two decoders of the batch in turn, each
    # on its own stream; a decoder's next stage (host planning + DMA + device
    # unstuffing, ordered on its stream) runs while the other decodes: files
    # in host memory to frames in HBM, PCIe included
    t_pipe = None
    if args.steps >= 2:
        pdecs = [batched_decoder() for _ in range(2)]
        pstreams = [torch.cuda.Stream() for _ in range(2)]
        for d, st in zip(pdecs, pstreams):  # buffers grown before the timed loop
            check(d.stage(files, st, asynchronous=True), "pipelined stage")
            d.decode(st)
            check(d.finish(st), "pipelined warm-up")
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for k in range(args.steps):
            d, st = pdecs[k % 2], pstreams[k % 2]
            if k >= 2:
                check(d.finish(st), "pipelined decode")
            check(d.stage(files, st, asynchronous=True), "pipelined stage")
            d.decode(st)
        for k in range(max(0, args.steps - 2), args.steps):
            check(pdecs[k % 2].finish(pstreams[k % 2]), "pipelined decode")
        torch.cuda.synchronize()
        t_pipe = (time.perf_counter() - t_start) / args.steps
        for d in pdecs:
            d.close()
        log(f"staging pipelined with decoding: {t_pipe * 1e3:.2f} ms per batch "
            f"({total_px / t_pipe / 1e9:.1f} Gpix/s, files in host memory to frames in HBM)")
    # frames back to host memory (the reader API's pixels buffer)
    t1 = time.perf_counter()
    for i in range(len(files)):
        b.pixels(i)
    d2h_ms = (time.perf_counter() - t1) * 1e3
    ms_ent, ms_pix = float(np.mean(ent)), float(np.mean(pix))
    ms_kent, ms_kexp = float(np.mean(kent)), float(np.mean(kexp))
    log(f"batched: {t_batched * 1e3:.2f} ms/step with {n_fl} in flight ({t_single * 1e3:.2f} ms with one), "
        f"{batched_mpix:.0f} Mpix/s; entropy stage {ms_ent:.2f} ms "
        f"(k_entropy {ms_kent:.2f}), pixel {ms_pix:.2f} ms")

    # algorithmic bytes per launch: the entropy decode reads the scan once and
    # yields 128 B of coefficients per block (k_entropy encodes them as
    # entries and block records); the pixel kernel reads the coefficients
    # (from those records) and writes the frames
    ent_bytes = tot["scan_bytes"] + tot["blocks"] * 128
    pix_bytes = tot["blocks"] * 128 + tot["out_bytes"]
    if ms_kent >= ms_pix:
        dom, dom_bytes, dom_ms = "cjd::k_entropy", ent_bytes, ms_kent
    else:
        dom, dom_bytes, dom_ms = f"cjd::k_pixel<{idct}>", pix_bytes, ms_pix
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    b.close()

    # ------------------------------------------------------------------ plain
    plain_mpix = None
    lat = []
    tune_table = None
    if not args.no_plain:
        # per-image entropy parameters for one-image launches, autotuned per
        # scan-size bucket on a few images of the batch
        from cudajpeg_amd.autotune import autotune
        t_at = time.time()
        tune_table = autotune(files, "latency", IDCT_FLOAT, reps=2, per_bucket=2)
        log(f"autotune (latency, {time.time() - t_at:.1f}s): {tune_table}")
        per = []
        for f in files:
            d = BatchDecoder(IDCT_FLOAT)  # entropy parameters from the tune table
            d.set_tune_table(tune_table)
            assert d.stage([f])[0] == 0
            per.append(d)
        with torch.cuda.stream(stream):
            for d in per:  # warm-up and latency: one image at a time
                d.decode(stream)
                assert d.finish(stream)[0] == 0
            for d in per:
                t1 = time.perf_counter()
                d.decode(stream)
                stream.synchronize()
                lat.append((time.perf_counter() - t1) * 1e3)
            torch.cuda.synchronize()
            t_start = time.perf_counter()
            for _ in range(args.steps):
                for d in per:
                    d.decode(stream)
            stream.synchronize()
            torch.cuda.synchronize()
            t_plain = (time.perf_counter() - t_start) / args.steps
        plain_mpix = total_px / t_plain / 1e6
        log(f"plain: {t_plain * 1e3:.2f} ms/step, {plain_mpix:.0f} Mpix/s")
        for d in per:
            d.close()

    progressive = None
    if not args.no_progressive:
        try:
            progressive = progressive_bench(args)
        except ImportError as e:  # Pillow writes the progressive files
            log(f"progressive files skipped: {e}")

    cpu = None if args.no_cpu else cpu_baseline(encs, args.cpu_seconds)
    if cpu is not None and args.cpu_workers > 1:
        cpu["all_cores"] = cpu_all_cores(encs, args.cpu_workers, 4)
        if cpu["all_cores"]:
            log(f"cpu baseline, {cpu['all_cores']['cores']} processes: {cpu['all_cores']['value']} Mpix/s")
    if cpu:
        log(f"cpu baseline: {cpu}")

    serving = None if args.no_serve else serve_child(args)
    if serving:
        log(f"serving: {serving}")

    out = {
        "metric": "Mpix/s batched-MFMA vs plain per-image on 1 MI355X; p50/p99 decode latency",
        "value": round(batched_mpix, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched * 1e3, 3),
        "ms_per_step_one_in_flight": round(t_single * 1e3, 3),
        "ms_per_step_graph": None if t_graph is None else round(t_graph * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"int": "int32", "float": "fp32", "mfma": "fp16x2+fp32acc"}[args.batched_idct],
        "data": "synthetic (seeded sinusoid+noise images, project baseline encoder)",
        "config": {"workload": f"mixed batch of {args.images} baseline JPEGs, 256x256..4096x3072",
                   "images": args.images, "mpix": round(total_px / 1e6, 1),
                   "scan_mb": round(tot["scan_bytes"] / 1e6, 1), "blocks": tot["blocks"],
                   "batched_idct": args.batched_idct, "plain_idct": "float", "parallelism": "1 GPU",
                   "launch": "direct", "parts": args.parts, "in_flight": n_fl,
                   "entropy_params": args.entropy_params},
        "pcie_inclusive": {"stage_ms": round(stage_ms, 2), "stage_ms_host_unstuff": round(stage_ms_host, 2),
                           "ms_per_batch_staging_pipelined": None if t_pipe is None else round(t_pipe * 1e3, 3),
                           "mpix_s_staging_pipelined": None if t_pipe is None else round(total_px / t_pipe / 1e6, 1),
                           "download_ms": round(d2h_ms, 2),
                           "mpix_s_with_stage": round(total_px / ((stage_ms + t_batched * 1e3) * 1e-3) / 1e6, 1),
                           "mpix_s_with_stage_and_download":
                               round(total_px / ((stage_ms + t_batched * 1e3 + d2h_ms) * 1e-3) / 1e6, 1)},
        "plain_mpix_s": None if plain_mpix is None else round(plain_mpix, 1),
        "plain_tune_table": tune_table,
        "ratio_batched_vs_plain": None if plain_mpix is None else round(batched_mpix / plain_mpix, 2),
        "p50_ms": None if not lat else round(float(np.percentile(lat, 50)), 3),
        "p99_ms": None if not lat else round(float(np.percentile(lat, 99)), 3),
        "kernel_ms": {"k_entropy": round(ms_kent, 3), "k_pixel": round(ms_pix, 3),
                      "entropy_stage": round(ms_ent, 3)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, args),
                     "algorithmic_bytes": int(dom_bytes)},
        "serving": serving,
        "progressive": progressive,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
