"""The product's kernel code (entropy_core.h, pixel_core.h) and host planner,
compiled for the CPU and run lane by lane, against the oracle: coefficient
parity of the tile/link/lookback entropy decoder and pixel parity of the fused
pixel stage, without a GPU."""
import sys

import numpy as np
import pytest

import oracle
from cudajpeg_amd import synth
from tests.emu import EmuResult

SPECS = [
    synth.ImageSpec(64, 48, "420", 75, 6.0, 1),
    synth.ImageSpec(33, 17, "444", 95, 8.0, 2),
    synth.ImageSpec(40, 24, "gray", 50, 8.0, 3),
    synth.ImageSpec(37, 29, "422", 80, 5.0, 7),
    synth.ImageSpec(100, 60, "420", 70, 6.0, 9, restart_interval=3),
    synth.ImageSpec(300, 200, "420", 95, 12.0, 11, random_tile_frac=0.3),
    synth.ImageSpec(512, 512, "420", 75, 6.0, 1000),
    synth.ImageSpec(640, 480, "444", 95, 8.0, 15),
    synth.ImageSpec(17, 9, "420", 90, 4.0, 12),
    synth.ImageSpec(200, 136, "420", 80, 6.0, 31, qtables=3),
    synth.ImageSpec(72, 40, "444", 60, 6.0, 32, qtables=3),
    # luma sampled 4x1, 1x2 and 4x2 (6, 4 and 10 blocks per MCU)
    synth.ImageSpec(131, 77, "411", 85, 6.0, 41),
    synth.ImageSpec(90, 101, "440", 75, 6.0, 42),
    synth.ImageSpec(150, 66, "410", 90, 8.0, 43, restart_interval=4),
]


@pytest.fixture(scope="module")
def encoded():
    return [synth.encode(s, truth=True) for s in SPECS]


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("tile_bits", [2048, 4096])
def test_emulated_kernels_bit_exact(encoded, order, tile_bits):
    r = EmuResult([e.data for e in encoded], 0, tile_bits, True, order)
    assert not r.errors.any()
    assert r.fallbacks == 0
    for i, (s, e) in enumerate(zip(SPECS, encoded)):
        p = oracle.parse(e.data)
        assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), e.coefs), s
        pl = oracle.planes(p, e.coefs, 0)
        for c in range(p.num_channels):
            assert np.array_equal(r.get(i, 2 + c, pl[c].shape, np.uint8), pl[c]), (s, c)
        ref = oracle.colour(p, pl)
        assert np.array_equal(r.get(i, 0, ref.shape, np.uint8), ref), s


def test_emulated_float_idct_within_one(encoded):
    r = EmuResult([e.data for e in encoded], 1, 4096, True)
    for i, (s, e) in enumerate(zip(SPECS, encoded)):
        p = oracle.parse(e.data)
        pl = oracle.planes(p, e.coefs, 1)
        for c in range(p.num_channels):
            got = r.get(i, 2 + c, pl[c].shape, np.uint8).astype(int)
            assert np.abs(got - pl[c]).max() <= 1, s


def test_emulated_serial_decoder(encoded):
    r = EmuResult([e.data for e in encoded], 0, 4096, False, 0, serial=True)
    for i, (s, e) in enumerate(zip(SPECS, encoded)):
        p = oracle.parse(e.data)
        assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), e.coefs), s


def test_emulated_large_high_entropy():
    # 4:2:0 q90 with 10% random tiles: long blocks stress tile links; the
    # lookback must still resolve without the serial fallback
    spec = synth.ImageSpec(1024, 576, "420", 90, 12.0, 2001, random_tile_frac=0.10)
    e = synth.encode(spec, truth=True)
    r = EmuResult([e.data], 0, 2048, False, 1)
    p = oracle.parse(e.data)
    assert r.fallbacks == 0
    assert np.array_equal(r.get(0, 1, (p.num_blocks, 64), np.int16), e.coefs)


HARD = [synth.ImageSpec(1200, 900, "420", 63, 8.9, 5),
        synth.ImageSpec(640, 480, "444", 95, 12.0, 15, random_tile_frac=0.5),
        synth.ImageSpec(800, 600, "gray", 90, 12.0, 3),
        synth.ImageSpec(1000, 700, "422", 80, 6.0, 9, restart_interval=7)]


@pytest.mark.parametrize("item_tiles,tile_bits,ck_bits", [(2, 2048, 0), (3, 4096, 0), (8, 4096, 128),
                                                          (8, 8192, 0)])
@pytest.mark.parametrize("order", [0, 1])
def test_emulated_small_items_chain_fixups(item_tiles, tile_bits, ck_bits, order):
    # Items of 2-8 tiles: many items per segment, so the lookback composes
    # chains of aggregates (order 1: every predecessor has only its
    # aggregate), and links / entries that miss the main trajectory take the
    # sequential fix-up.  Both must occur here and the result stay bit-exact.
    encs = [synth.encode(s, truth=True) for s in HARD]
    r = EmuResult([e.data for e in encs], 0, tile_bits, False, order, False, item_tiles, ck_bits)
    assert not r.errors.any() and r.fallbacks == 0
    assert r.fixes > 0 and r.items > 4 * len(HARD)
    for i, (s, e) in enumerate(zip(HARD, encs)):
        assert np.array_equal(r.get(i, 1, e.coefs.shape, np.int16), e.coefs), s


@pytest.mark.parametrize("seed", [3, 17, 91, 1234])
def test_emulated_lookback_in_random_orders(seed):
    """The lookback resolved in a random order of the work items (emu order
    >= 3, seeded): every walk meets its own mix of inclusive and
    aggregate-only predecessors, as uneven progress of the workgroups on the
    GPU gives it; 3-tile items of 512-bit tiles make chains tens of items
    long.  Coefficients stay exactly the encoder's."""
    specs = [s for s in synth.config_mixed(24, 777) if s.pixels < 3_000_000][:8]
    encs = [synth.encode(s, truth=True) for s in specs]
    r = EmuResult([e.data for e in encs], 0, 512, False, seed, False, 3)
    assert r.items > 1000
    for i, (s, e) in enumerate(zip(specs, encs)):
        assert r.errors[i] == 0, s
        assert np.array_equal(r.get(i, 1, e.coefs.shape, np.int16), e.coefs), s


def test_emulated_corrupt_scan_reports_syntax():
    enc = synth.encode(synth.ImageSpec(64, 64, "420", 75, 6.0, 21))
    data = bytearray(enc.data)
    p = oracle.parse(bytes(data))
    start = len(data) - 2 - p._raw.scan_len
    mid = start + p._raw.scan_len // 2
    data[mid:mid + 6] = b"\xff\x00\xff\x00\xff\x00"
    assert oracle.decode_scan(oracle.parse(bytes(data))).error == 1
    r = EmuResult([bytes(data)])
    assert r.errors[0] == 1


def test_emulated_mfma_contraction_matches_fp64(encoded):
    # host restatement of the MFMA iDCT (FP16 hi/lo basis, FP32 sums) is within
    # 1 of the FP64 iDCT and within 1 of the reference's integer iDCT
    r = EmuResult([e.data for e in encoded], 2, 4096, True)
    for i, (s, e) in enumerate(zip(SPECS, encoded)):
        p = oracle.parse(e.data)
        f64 = oracle.planes(p, e.coefs, 1)
        i32 = oracle.planes(p, e.coefs, 0)
        for c in range(p.num_channels):
            got = r.get(i, 2 + c, f64[c].shape, np.uint8).astype(int)
            assert np.abs(got - f64[c]).max() <= 1, s
            assert (got == f64[c]).mean() > 0.999, s
            assert np.abs(got - i32[c]).max() <= 1, s


@pytest.mark.parametrize("tile_bits,item_tiles", [(1024, 0), (4096, 0), (512, 3)])
def test_emulated_long_code_tables(tile_bits, item_tiles):
    # AC tables with 60 codes of 11 bits and 92 of 16: the image's long codes
    # need more second-level lookup tables than the pool holds, so the chroma
    # AC table decodes its long codes canonically; parity must not change.
    # Items of 3 tiles: entries differ from the assumed ones, so links run
    # again after the counts (which overlay the canonical tables in LDS).
    specs = [synth.ImageSpec(320, 240, "444", 90, 8.0, 71, huff=1),
             synth.ImageSpec(400, 300, "420", 75, 8.0, 72, huff=1, restart_interval=5)]
    encs = [synth.encode(s, truth=True) for s in specs]
    r = EmuResult([e.data for e in encs], 0, tile_bits, False, 0, False, item_tiles)
    assert not r.errors.any()
    if item_tiles:
        assert r.fixes > 0
    for i, (s, e) in enumerate(zip(specs, encs)):
        assert np.array_equal(r.get(i, 1, e.coefs.shape, np.int16), e.coefs), s



def test_emulated_lookback_give_up_is_redone_serially():
    # Every wait gives up at once and segment-first items resolve last (order
    # 2): every work item after the first of its segment fails its lookback,
    # publishes the failed state, and its image is marked unresolved; the host
    # decodes those images again with the serial decoder, and the results stay
    # bit-exact.  Images with one work item per segment are not affected.
    encs = [synth.encode(s, truth=True) for s in HARD]
    r = EmuResult([e.data for e in encs], 0, 2048, False, 2, False, 4)
    assert not r.errors.any()
    # every image without restart markers has segments of several items (the
    # restart-interval image's segments of 7 MCUs are one item each)
    assert r.fallbacks == sum(1 for s in HARD if not s.restart_interval)
    for i, (s, e) in enumerate(zip(HARD, encs)):
        assert np.array_equal(r.get(i, 1, e.coefs.shape, np.int16), e.coefs), s
    small = synth.encode(synth.ImageSpec(64, 48, "420", 75, 6.0, 1), truth=True)
    r = EmuResult([small.data], 0, 4096, False, 2)
    assert r.fallbacks == 0 and np.array_equal(r.get(0, 1, small.coefs.shape, np.int16), small.coefs)


def test_emulated_inconsistent_checkpoints_stay_inside_their_image():
    # Failure injection on the CPU: checkpoint words that claim more DC
    # entries than their speculative pass decoded (as a word this pass never
    # wrote would).  A link that joins at one of them must not produce a
    # negative block count -- which would make the record phase write block
    # records below its image's first block, into other images' records or
    # outside the buffer -- but must leave the image to the serial decoder:
    # every image of the batch still decodes bit-exact.
    import ctypes
    from tests.emu import lib
    L = lib()
    L.emu_plant_ck.argtypes = [ctypes.c_int]
    encs = [synth.encode(s, truth=True) for s in HARD]
    L.emu_plant_ck(5)
    try:
        r = EmuResult([e.data for e in encs], 0, 2048, False, 0, False, 8)
    finally:
        L.emu_plant_ck(0)
    assert not r.errors.any()
    assert r.fallbacks > 0
    for i, (s, e) in enumerate(zip(HARD, encs)):
        assert np.array_equal(r.get(i, 1, e.coefs.shape, np.int16), e.coefs), s


_TSAN_CHILD = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cudajpeg_amd import synth
L = ctypes.CDLL(sys.argv[2])
L.emt_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_void_p]
specs = [synth.ImageSpec(487, 487, "420", 88, 9.0, 77), synth.ImageSpec(400, 300, "444", 95, 12.0, 15)]
encs = [synth.encode(s, truth=True) for s in specs]
n = len(encs)
keep = [ctypes.create_string_buffer(e.data, len(e.data)) for e in encs]
ptrs = (ctypes.c_void_p * n)(*[ctypes.cast(b, ctypes.c_void_p) for b in keep])
sizes = (ctypes.c_size_t * n)(*[len(e.data) for e in encs])
for tile, items in ((2048, 8), (4096, 256)):
    outs = [np.zeros_like(e.coefs) for e in encs]
    optr = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    bits = np.zeros(n, np.uint32)
    L.emt_run(n, ptrs, sizes, tile, items, optr, bits.ctypes.data)
    for e, o, b in zip(encs, outs, bits):
        print("exact", int(b == 0 and np.array_equal(o, e.coefs)), flush=True)
"""


def test_kernel_body_has_no_lane_races_under_thread_sanitizer():
    # The body of k_entropy run with one host thread per lane and a barrier
    # wherever the kernel has one (tests/emu/emu_threads.cpp), built with
    # ThreadSanitizer: lanes really run at the same time between barriers, so
    # an unsynchronised conflicting access between lanes (LDS state or entry
    # buffer) is reported.  Items of 8 tiles (many links, fix-ups and
    # lookbacks) and of 256 tiles; results must be exact as well.
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
    so = os.path.join(here, "libcjd_emt_tsan.so")
    subprocess.run(["make", "-s", "-C", here], check=True)
    rt = subprocess.run(["g++", "-print-file-name=libtsan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isfile(rt):
        pytest.skip("no ThreadSanitizer runtime")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = {**os.environ, "LD_PRELOAD": rt, "TSAN_OPTIONS": "report_signal_unsafe=0 halt_on_error=0 exitcode=0"}
    r = subprocess.run([sys.executable, "-c", _TSAN_CHILD, root, so], capture_output=True, text=True, env=env,
                       timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.count("exact 1") == 4, r.stdout


def test_emulated_device_unstuff_equals_host_on_files(encoded):
    # the device unstuffer (unstuff_core.h, emulated lane by lane) writes the
    # batch's word buffer exactly as the host planner does, padding included,
    # for every sampling, restart intervals and random tiles, and files large
    # enough for many chunks
    from tests.emu import unstuff_files
    files = [e.data for e in encoded] + [synth.encode(synth.ImageSpec(1200, 900, "420", 95, 12.0, 77)).data,
                                         synth.encode(synth.ImageSpec(700, 500, "422", 90, 9.0, 78,
                                                                      restart_interval=2)).data]
    r = unstuff_files(files)
    assert not isinstance(r, int), r
    host, dev = r
    assert len(host) > 4 * 4096 // 4
    assert np.array_equal(host, dev)


def _adversarial_scan(rng, n, split):
    """Random bytes with many 0xFF: pairs with 0x00, runs of 0xFF, other marker
    bytes, RSTn in sequence (split) or anywhere, some at chunk boundaries."""
    out = bytearray()
    rst = 0
    while len(out) < n:
        r = rng.random()
        if r < 0.55:
            out += bytes(rng.integers(0, 255, int(rng.integers(1, 64)), dtype=np.uint8))  # no 0xFF
        elif r < 0.75:
            out += b"\xff\x00"
        elif r < 0.85:
            out += b"\xff" * int(rng.integers(1, 9))
        elif r < 0.92:
            out += bytes([0xFF, 0xD0 + rst])
            rst = (rst + 1) & 7
        else:
            out += bytes([0xFF, int(rng.integers(1, 0xCF))])
        if rng.random() < 0.05:  # pad to just before / at / after a chunk boundary
            k = 4096 - len(out) % 4096 + int(rng.integers(-2, 3))
            out += b"\x11" * max(0, k)
    return bytes(out[:n])


@pytest.mark.parametrize("split", [False, True])
def test_emulated_device_unstuff_equals_host_on_adversarial_bytes(split):
    from tests.emu import unstuff_bytes
    rng = np.random.default_rng(9100 + split)
    done = 0
    for case in range(60):
        n = int(rng.choice([1, 2, 15, 16, 17, 4095, 4096, 4097, 8191, 8192, 8193, int(rng.integers(1, 40000))]))
        data = _adversarial_scan(rng, n, split)
        if case % 7 == 3:
            data = data[:-1] + b"\xff"  # a 0xFF as the scan's last byte
        r = unstuff_bytes(data, split)
        if r == -3:  # out-of-sequence RSTn after truncation: both refuse
            continue
        assert not isinstance(r, int), (case, n, r)
        host, dev = r
        assert np.array_equal(host, dev), (case, n, int(np.flatnonzero(host != dev)[0]))
        done += 1
    assert done >= 40


DQT16_SPECS = [synth.ImageSpec(96, 64, "gray", 3, 6.0, 91, dqt16=1),
               synth.ImageSpec(130, 70, "420", 5, 8.0, 92, dqt16=1, qtables=3),
               synth.ImageSpec(64, 64, "444", 8, 10.0, 93, dqt16=1, restart_interval=5)]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_emulated_dqt16_extension(mode):
    # 16-bit quant tables with CJD_EXT_DQT16: coefficients from the encoder,
    # integer-iDCT planes and pixels equal the oracle's (with the extension),
    # FP32 and MFMA planes within 1 of its FP64 planes; without the extension
    # every such file is refused with SYNTAX_ERROR, as by the reference
    from tests.emu import lib as emu_lib
    encs = [synth.encode(s, truth=True) for s in DQT16_SPECS]
    r0 = EmuResult([e.data for e in encs], mode, 4096, True)
    assert (r0.errors == 1).all()
    emu_lib().emu_ext(1)
    try:
        r = EmuResult([e.data for e in encs], mode, 4096, True)
    finally:
        emu_lib().emu_ext(0)
    assert not r.errors.any()
    for i, (s, e) in enumerate(zip(DQT16_SPECS, encs)):
        p = oracle.parse(e.data, oracle.EXT_DQT16)
        assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), e.coefs), s
        pl = oracle.planes(p, e.coefs, 0 if mode == 0 else 1)
        for c in range(p.num_channels):
            got = r.get(i, 2 + c, pl[c].shape, np.uint8).astype(int)
            if mode == 0:
                assert np.array_equal(got, pl[c]), (s, c)
            else:
                assert np.abs(got - pl[c]).max() <= 1, (s, c)
        if mode == 0:
            assert np.array_equal(r.get(i, 0, (s.height, s.width, s.ncomp), np.uint8), oracle.colour(p, pl)), s
