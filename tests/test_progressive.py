"""Progressive JPEG (SOF2, T.81 G.1) behind CJD_EXT_PROGRESSIVE
(ORC_EXT_PROGRESSIVE in the oracle).  The reference refuses SOF2 as an
unknown marker (src/format.cu:198-200) with SYNTAX_ERROR, and both decoders
do the same without the flag.

The oracle's progressive decode is pinned two ways:
* by libjpeg's baseline and progressive files of the same pixels
  (tests/golden/progressive.npz, written by make_progressive.py): the
  baseline decode is pinned by the reference, and the two files hold the same
  quantised coefficients, so the progressive decode must give exactly those
  coefficients and pixels;
* scan by scan: the file cut after its n-th scan (an EOI appended) is a
  valid progressive file whose coefficients are the final ones under T.81's
  point transform of the bits coded so far -- a DC value v coded down to bit
  Al reads (v >> Al) << Al, an AC value sign(v) (|v| >> Al) << Al, an uncoded
  one 0.  Every scan is checked on its own: a wrong end-of-band run, a
  missed correction bit or a misplaced refinement coefficient changes the
  prefix that ends with that scan.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from cudajpeg_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
PR = oracle.EXT_PROGRESSIVE
_Z = np.load(os.path.join(HERE, "golden", "progressive.npz"))
NAMES = sorted({k.split("/")[0] for k in _Z.files})


def pair(name):
    return _Z[name + "/base"].tobytes(), _Z[name + "/prog"].tobytes()


def scans(p):
    return list(p._raw.scans)[: p._raw.num_scans]


def scan_bounds(p, s):
    base = ctypes.cast(ctypes.c_char_p(p._data), ctypes.c_void_p).value
    return s.scan - base, s.scan - base + s.scan_len


def slot_comps(p):
    out = []
    for c, ch in enumerate(p.channels):
        out += [c] * (ch["samples_x"] * ch["samples_y"])
    return np.array(out)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_pinned_by_baseline_twin(name):
    base, prog = pair(name)
    assert oracle.parse(prog).error == 1  # the reference's SYNTAX_ERROR without the flag
    pb, pp = oracle.parse(base), oracle.parse(prog, PR)
    assert pb.error == 0 and pp.error == 0
    assert pp._raw.progressive == 1 and pp._raw.num_scans >= 6
    rb, rp = oracle.decode_scan(pb), oracle.decode_scan(pp)
    assert rb.error == 0 and rp.error == 0
    assert np.array_equal(rp.qcoef, rb.qcoef)
    assert np.abs(rb.qcoef).max() > 0
    e1, x1 = oracle.decode(base)
    e2, x2 = oracle.decode(prog, 0, PR)
    assert e1 == e2 == 0 and np.array_equal(x1, x2)
    # the dequantised working space, as the baseline decode lays it out
    wb = oracle.decode_scan(pb, want_work=True).work
    wp = oracle.decode_scan(pp, want_work=True).work
    for a, b in zip(wb, wp):
        assert np.array_equal(a, b)


def point_transform(final, bits, dc):
    """Coefficients coded down to bit `bits` (-1: not coded)."""
    coded = np.broadcast_to(bits >= 0, final.shape)
    sh = np.broadcast_to(np.where(bits >= 0, bits, 0), final.shape).astype(np.int32)
    v = final.astype(np.int32)
    if dc:
        t = (v >> sh) << sh
    else:
        t = np.sign(v) * ((np.abs(v) >> sh) << sh)
    return np.where(coded, t, 0).astype(final.dtype)


ZZ = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20,
               13, 6, 7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52,
               45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_scan_prefixes_follow_point_transform(name):
    base, prog = pair(name)
    final = oracle.decode_scan(oracle.parse(base)).qcoef
    pp = oracle.parse(prog, PR)
    sc = scans(pp)
    comp_of_block = np.tile(slot_comps(pp), pp.mcus_x * pp.mcus_y)
    bits = -np.ones((pp.num_channels, 64), int)  # per component, zigzag index
    changed = 0
    prev = np.zeros_like(final)
    for n, s in enumerate(sc):
        for i in range(s.ncomp):
            bits[s.comps[i], s.ss: s.se + 1] = s.al
        _, end = scan_bounds(pp, s)
        cut = prog[:end] + b"\xff\xd9"
        pc = oracle.parse(cut, PR)
        assert pc.error == 0 and pc._raw.num_scans == n + 1
        r = oracle.decode_scan(pc)
        assert r.error == 0, n
        want = np.zeros_like(final)
        for c in range(pp.num_channels):
            blk = comp_of_block == c
            nat = np.zeros(64, int)
            nat[ZZ] = bits[c]
            want[blk, :1] = point_transform(final[blk, :1], nat[None, :1], True)
            want[blk, 1:] = point_transform(final[blk, 1:], nat[None, 1:], False)
        assert np.array_equal(r.qcoef, want), (name, n, (s.ncomp, s.ss, s.se, s.ah, s.al))
        changed += not np.array_equal(r.qcoef, prev)
        prev = r.qcoef
    if pp.num_blocks >= 200:  # every scan of the script changes some coefficient of a larger image
        assert changed == len(sc)
    assert np.array_equal(prev, final)


def sos_headers(data):
    out, k = [], data.find(b"\xff\xda")
    while k >= 0:
        out.append(k)
        k = data.find(b"\xff\xda", k + 2)
    return out


def _parse_product(data, ext):
    info = L.FrameInfo()
    e = L.lib().cjd_parse_ext(data, len(data), ext, ctypes.byref(info))
    return e, info


def malformed_cases():
    """Progressions T.81 G.1.1.1 forbids, each from a valid 4:2:0 file of the
    standard script (scan 0: DC first Al 1; 1: Y 1-5 Al 2; 4: Y 6-63 Al 2;
    5: Y 1-63 Ah 2 Al 1)."""
    _, prog = pair("c420_64x48_q75")
    pp = oracle.parse(prog, PR)
    sos = sos_headers(prog)
    assert len(sos) == 10

    def setb(data, off, val):
        return data[:off] + bytes([val]) + data[off + 1:]

    def ss_at(k):  # the Ss byte of scan k's SOS header
        ns = prog[sos[k] + 4]
        return sos[k] + 5 + 2 * ns

    s0_end = scan_bounds(pp, scans(pp)[0])[1]
    return {
        "refinement by two bits": setb(prog, ss_at(5) + 2, 0x20),
        "Ah not the band's last Al": setb(prog, ss_at(5) + 2, 0x32),
        "DC scan with Se > 0": setb(prog, ss_at(1), 0),
        "band coded twice as first": setb(prog, ss_at(4), 5),
        "Se past 63": setb(prog, ss_at(4) + 1, 64),
        "Se before Ss": setb(prog, ss_at(4) + 1, 5),
        "Al past 13": setb(prog, ss_at(1) + 2, 14),
        "AC before the component's DC": prog[: sos[0]] + prog[s0_end:],
        "unknown component": setb(prog, sos[1] + 5, 9),
    }


@pytest.mark.parametrize("case", sorted(malformed_cases()))
def test_malformed_progression(case):
    data = malformed_cases()[case]
    assert oracle.parse(data, PR).error == 1, case
    assert _parse_product(data, L.EXT_PROGRESSIVE)[0] == 1, case


@pytest.mark.parametrize("name", NAMES)
def test_product_parser(name):
    _, prog = pair(name)
    assert _parse_product(prog, 0)[0] == 1  # the reference's SYNTAX_ERROR
    e, info = _parse_product(prog, L.EXT_PROGRESSIVE)
    p = oracle.parse(prog, PR)
    assert e == 0
    assert info.num_scans == p._raw.num_scans
    assert (info.width, info.height, info.num_channels) == (p.width, p.height, p.num_channels)
    assert info.progressive == 1
    # each component's quant table is the one in effect at its first scan
    for c in range(p.num_channels):
        assert list(info.dq_tables16[info.dq_id[c]]) == list(p._raw.comp_dq16[c])


def pillow_progressive(w, h, seed, gray=False, q=85, **opts):
    """A progressive file written by libjpeg through Pillow (imported here
    only: the tests that need larger files than the committed fixtures)."""
    Image = pytest.importorskip("PIL.Image")
    import io
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_progressive", os.path.join(HERE, "golden",
                                                                                    "make_progressive.py"))
    mp = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mp)
    b = io.BytesIO()
    Image.fromarray(mp.pixels(w, h, seed, gray)).save(b, "JPEG", quality=q, progressive=True, **opts)
    return b.getvalue()


@pytest.mark.parametrize("serial", [False, True])
def test_emulator_matches_oracle(serial):
    """The planner (scan levels, restart intervals as tasks), k_prog_prepare
    and k_prog (prog_core.h), lane by lane on the CPU, beside baseline files
    in the same batch: coefficients equal the oracle's, integer-iDCT pixels
    and planes too."""
    from tests.emu import EmuResult
    from tests.emu import lib as emulib
    from cudajpeg_amd import synth
    files = [pair(n)[1] for n in NAMES]
    files += [pillow_progressive(203, 150, 21, restart_marker_rows=2),
              pillow_progressive(120, 64, 22, gray=True, q=95),
              pillow_progressive(640, 480, 24, q=95)]  # scans of many stream refills, blocks of many batches
    base = [synth.encode(synth.ImageSpec(96, 64, "420", 80, 6.0, 23)).data]
    files = files[:3] + base + files[3:]
    emulib().emu_ext(L.EXT_PROGRESSIVE)
    try:
        r = EmuResult(files, 0, 512, True, 0, serial)
    finally:
        emulib().emu_ext(0)
    for i, data in enumerate(files):
        assert r.errors[i] == 0, i
        p = oracle.parse(data, PR)
        q = oracle.decode_scan(p).qcoef
        assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), q), i
        pl = oracle.planes(p, q, 0)
        for c in range(p.num_channels):
            assert np.array_equal(r.get(i, 2 + c, pl[c].shape, np.uint8), pl[c]), (i, c)
        ref = oracle.colour(p, pl)
        assert np.array_equal(r.get(i, 0, ref.shape, np.uint8), ref), i


def test_emulator_matches_oracle_on_corrupt_streams():
    """Bytes of one scan's data overwritten at random (never making or
    breaking a marker): the product's decoder (emulated) and the oracle give
    the same verdict, and the same coefficients when both decode the file.
    Between restart intervals both require the interval's data to end in its
    last byte before the RSTn; data after a scan's last interval is
    ignored."""
    from tests.emu import EmuResult
    from tests.emu import lib as emulib
    rs = np.random.RandomState(5)
    files = []
    for it in range(330):
        prog = bytearray(pair(NAMES[it % len(NAMES)])[1])
        pp = oracle.parse(bytes(prog), PR)
        sc = scans(pp)
        a, b = scan_bounds(pp, sc[rs.randint(len(sc))])
        for _ in range(rs.randint(1, 4)):
            pos = rs.randint(a, b) if b > a else a
            if pos >= b or prog[pos] == 0xFF or prog[pos - 1] == 0xFF:
                continue
            prog[pos] = min(rs.randint(256), 0xFE)
        files.append(bytes(prog))
    emulib().emu_ext(L.EXT_PROGRESSIVE)
    try:
        r = EmuResult(files, 0, 512, True, 0, False)
    finally:
        emulib().emu_ext(0)
    verdicts = {0: 0, 1: 0}
    for i, d in enumerate(files):
        p = oracle.parse(d, PR)
        assert p.error == 0
        o = oracle.decode_scan(p)
        assert r.errors[i] == o.error, i
        verdicts[o.error] += 1
        if not o.error:
            assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), o.qcoef), i
    assert verdicts[0] > 30 and verdicts[1] > 30  # both outcomes are exercised
