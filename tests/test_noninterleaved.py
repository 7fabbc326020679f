"""Non-interleaved baseline scans: one scan per component (T.81 A.2.2), behind
CJD_EXT_NONINTERLEAVED (ORC_EXT_NONINTERLEAVED in the oracle).  The
reference refuses them: its SOS check (src/decodeScanGPU.cu:124-131) fails a
one-component scan of a three-component frame on the header length first,
with SYNTAX_ERROR, and the product does the same without the flag.

The oracle's decode of such files is pinned by the same images encoded as
one interleaved scan, whose decode the reference pins.  The pixels are
identical: the blocks by which MCUs pad a component's own grid are not coded
in a one-component scan, and they lie outside the image.  The coefficients
equal the encoder's at every coded block and are zero at the padding.  The
encoder redefines Huffman table 0 with the chroma tables between the luma and
chroma scans, and with restarts changes the interval there, so a decoder that
took the file's last tables or interval for every scan fails.  Its second
form also redefines quant table 1 before the Cr scan.  That quantises Cr as
the three-table interleaved file does, so equal pixels also pin that each
component takes the quant table in effect when its scan begins.
"""
import ctypes

import numpy as np
import pytest

import oracle
from cudajpeg_amd import _lib as L
from cudajpeg_amd import synth
from tests.emu import EmuResult
from tests.emu import lib as emulib

NI = oracle.EXT_NONINTERLEAVED
SAMPLINGS = ["420", "422", "444", "411", "440", "410"]
SIZES = [(333, 217), (17, 9), (1001, 77)]


def _pair(sampling, w, h, ri, mode, seed):
    """(interleaved, non-interleaved) encodings with the same quantisation."""
    inter = synth.ImageSpec(w, h, sampling, 80, 6.0, seed, restart_interval=ri, qtables=3 if mode == 2 else 2)
    non = synth.ImageSpec(w, h, sampling, 80, 6.0, seed, restart_interval=ri, noninterleaved=mode)
    return synth.encode(inter, truth=True), synth.encode(non, truth=True)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("ri", [0, 3])
@pytest.mark.parametrize("sampling", SAMPLINGS)
def test_oracle_pinned_by_interleaved_encoding(sampling, ri, mode):
    for k, (w, h) in enumerate(SIZES):
        ei, en = _pair(sampling, w, h, ri, mode, 500 + k)
        e1, pi = oracle.decode(ei.data)
        e2, pn = oracle.decode(en.data, 0, NI)
        assert e1 == 0 and e2 == 0, (sampling, w, h)
        assert np.array_equal(pi, pn), (sampling, w, h, ri, mode)
        p = oracle.parse(en.data, NI)
        r = oracle.decode_scan(p)
        assert r.error == 0
        assert np.array_equal(r.qcoef, en.coefs), (sampling, w, h)
        # coded blocks as in the interleaved file, the padding zero
        coded = _coded(p)
        assert coded.sum() == sum(((c["width"] + 7) // 8) * ((c["height"] + 7) // 8) for c in p.channels)
        assert np.array_equal(en.coefs[coded], ei.coefs[coded])
        assert not en.coefs[~coded].any()


def _coded(p):
    """Blocks of the MCU layout that a one-component scan codes: those of
    component c's own grid of ceil(width_c / 8) x ceil(height_c / 8)."""
    bpm = sum(c["samples_x"] * c["samples_y"] for c in p.channels)
    m = np.zeros((p.mcus_y, p.mcus_x, bpm), bool)
    slot0 = 0
    for c in p.channels:
        h, v = c["samples_x"], c["samples_y"]
        by, bx = np.mgrid[0:(c["height"] + 7) // 8, 0:(c["width"] + 7) // 8]
        m[by // v, bx // h, slot0 + (by % v) * h + bx % h] = True
        slot0 += h * v
    return m.reshape(-1)


def _parse_product(data, ext):
    info = L.FrameInfo()
    e = L.lib().cjd_parse_ext(data, len(data), ext, ctypes.byref(info))
    return e, info


def test_refused_without_extension_as_the_reference():
    _, en = _pair("420", 64, 48, 0, 1, 7)
    assert oracle.parse(en.data).error == oracle.parse(en.data, 0).error == 1  # SYNTAX_ERROR
    e, _ = _parse_product(en.data, 0)
    assert e == 1
    e, info = _parse_product(en.data, L.EXT_NONINTERLEAVED)
    assert e == 0 and info.num_scans == 3
    p = oracle.parse(en.data, NI)
    assert info.scan_offset == p._raw.scans[0].scan - ctypes.cast(ctypes.c_char_p(p._data), ctypes.c_void_p).value
    assert info.scan_length == p._raw.scans[0].scan_len
    # an interleaved file with the flag: one scan, as before
    ei, _ = _pair("420", 64, 48, 0, 1, 7)
    e, info = _parse_product(ei.data, L.EXT_NONINTERLEAVED)
    assert e == 0 and info.num_scans == 1


def _sos_positions(data):
    out, k = [], data.find(b"\xff\xda")
    while k >= 0:
        out.append(k)
        k = data.find(b"\xff\xda", k + 2)
    return out


def test_malformed_scan_sets():
    """A component without a scan, a component in two scans, or an interleaved
    scan after one-component scans: SYNTAX_ERROR from both parsers."""
    _, en = _pair("422", 100, 60, 0, 1, 9)
    sos = _sos_positions(en.data)
    assert len(sos) == 3
    cases = {
        "missing Cr": en.data[: sos[2]] + b"\xff\xd9",
        "Cb scan codes Y again": en.data[: sos[1] + 5] + b"\x01" + en.data[sos[1] + 6:],
        "interleaved scan last": en.data[: sos[2]] + bytes([0xFF, 0xDA, 0, 12, 3, 1, 0, 2, 0x11, 3, 0x11, 0, 63, 0])
        + b"\x00" * 8 + b"\xff\xd9",
    }
    for name, data in cases.items():
        assert oracle.parse(data, NI).error == 1, name
        assert _parse_product(data, L.EXT_NONINTERLEAVED)[0] == 1, name


@pytest.mark.parametrize("tile_bits,order,serial", [(512, 0, False), (4096, 0, False), (512, 2, False),
                                                    (1024, 0, True)])
def test_emulator_matches_oracle(tile_bits, order, serial):
    """The planner and the tiled decoder's phases (entropy_core.h), lane by
    lane on the CPU: coefficients equal the encoder's, integer-iDCT pixels
    and planes the oracle's, in tiles of 512 bits (many work items per scan
    segment, lookback across items), ordered launches and the serial
    decoder."""
    specs = [synth.ImageSpec(w, h, s, 80, 6.0, 300 + i, restart_interval=ri, noninterleaved=m)
             for i, s in enumerate(SAMPLINGS) for (w, h) in SIZES[:2] for ri in (0, 3) for m in (1, 2)]
    encs = [synth.encode(s, truth=True) for s in specs]
    emulib().emu_ext(L.EXT_NONINTERLEAVED)
    try:
        r = EmuResult([e.data for e in encs], 0, tile_bits, True, order, serial)
    finally:
        emulib().emu_ext(0)
    for i, (s, e) in enumerate(zip(specs, encs)):
        assert r.errors[i] == 0, s
        p = oracle.parse(e.data, NI)
        assert np.array_equal(r.get(i, 1, (p.num_blocks, 64), np.int16), e.coefs), s
        pl = oracle.planes(p, e.coefs, 0)
        for c in range(p.num_channels):
            assert np.array_equal(r.get(i, 2 + c, pl[c].shape, np.uint8), pl[c]), (s, c)
        ref = oracle.colour(p, pl)
        assert np.array_equal(r.get(i, 0, ref.shape, np.uint8), ref), s
