"""The oracle's layout, packetize and packet-parser restatements against the
golden fixtures written from the reference's own metal/pyrowave_bitstream.cpp
(tests/golden/make_golden.py).  CPU only."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LAYOUT_FILES = sorted(glob.glob(os.path.join(GOLDEN, "layout_*.npz")))


def _dims(path):
    name = os.path.basename(path)[len("layout_"):-len(".npz")]
    wh, ch = name.split("_")
    w, h = wh.split("x")
    return int(w), int(h), 0 if ch == "420" else 1


def test_fixtures_present():
    assert len(LAYOUT_FILES) >= 10
    for f in ("packetize.npz", "parser.npz"):
        assert os.path.exists(os.path.join(GOLDEN, f))


@pytest.mark.parametrize("path", LAYOUT_FILES, ids=[os.path.basename(p) for p in LAYOUT_FILES])
def test_layout_matches_reference(path):
    """Block counts, per-band offsets/strides, the 32x32 -> 8x8 mapping and the
    subdivision size (pyrowave_common.cpp:191-254, pyrowave_bitstream.cpp:45-95)."""
    g = np.load(path)
    w, h, ch = _dims(path)
    L = O.layout(w, h, ch)
    assert L.count8 == int(g["count8"])
    assert L.count32 == int(g["count32"])
    assert L.per_subdivision == int(g["per_subdivision"])
    band = g["band"]
    for c in range(3):
        for l in range(5):
            for b in range(4):
                B = L.band[c][l][b]
                off8, stride8, off32, stride32 = (int(v) for v in band[c, l, b])
                if not B.present:
                    continue
                assert (B.off8, B.bx8, B.off32, B.bx32) == (off8, stride8, off32, stride32), (c, l, b)
    # Mapping of every 32x32 block to its 8x8 footprint, from the oracle's
    # block descriptors.
    d = O.block_descs(L)
    mapping = g["mapping"]
    for i in range(L.count32):
        B = L.band[d["component"][i]][d["level"][i]][d["band"][i]]
        bx, by = int(d["bx"][i]), int(d["by"][i])
        exp = (B.off8 + 4 * by * B.bx8 + 4 * bx, B.bx8, min(4, B.bx8 - 4 * bx), min(4, B.by8 - 4 * by))
        assert tuple(int(v) for v in mapping[i]) == exp, i


def test_layout_block_order_is_level_major():
    """Block indices run level 4 -> 0, then component, then band
    (bitstream/bitstream.md:99-152): offsets strictly increase in that order."""
    L = O.layout(1920, 1080, 0)
    last = -1
    for l in range(4, -1, -1):
        for c in range(3):
            for b in range(4):
                B = L.band[c][l][b]
                if B.present:
                    assert B.off32 > last
                    last = B.off32


def _packetize_cases():
    g = np.load(os.path.join(GOLDEN, "packetize.npz"))
    return g, int(g["count"])


def test_packetize_matches_reference():
    """compute_num_packets and packetize (pyrowave_encoder.cpp:596-623,
    1048-1115) at several boundaries, including an output buffer too small."""
    g, n = _packetize_cases()
    assert n >= 20
    for i in range(n):
        j, boundary, out_size, npk = (int(v) for v in g[f"c{i}_params"])
        w, h, ch = (int(v) for v in g[f"l{j}_dims"])
        meta, bits = g[f"l{j}_meta"], g[f"l{j}_bits"]
        L = O.layout(w, h, ch)
        assert O.compute_num_packets(L, meta, boundary) == npk, i
        out, packets = O.packetize(L, meta, bits, 5, boundary, out_size=out_size)
        exp_packets = g[f"c{i}_packets"].astype(np.int64)
        assert packets.shape == exp_packets.shape, i
        assert np.array_equal(packets, exp_packets), i
        written = int(packets[-1].sum()) if len(packets) else 0
        assert np.array_equal(out[:written], g[f"c{i}_out"][:written]), i


def _parser():
    return np.load(os.path.join(GOLDEN, "parser.npz"))


PARSER_NAMES = [str(n) for n in _parser()["names"]]


@pytest.mark.parametrize("name", PARSER_NAMES)
def test_parser_matches_reference(name):
    """The decoder's packet parser (pyrowave_decoder.cpp:126-262) step by step:
    push result, ready (whole and partial), decoded and total blocks, payload
    size; then the final offsets and payload."""
    g = _parser()
    w, h, ch = int(g["w"]), int(g["h"]), int(g["chroma"])
    L = O.layout(w, h, ch)
    dec = O.Decoder(L)
    ops, lens, data = g[f"{name}_ops"], g[f"{name}_lens"], g[f"{name}_data"].tobytes()
    steps = g[f"{name}_steps"]
    pos = 0
    for k, (op, n) in enumerate(zip(ops, lens)):
        rc = -1
        if op == 0:
            rc = int(dec.push(data[pos:pos + n]))
            pos += int(n)
        elif op == 1:
            dec.clear()
        else:
            dec.mark_decoded()
        got = (rc, int(dec.ready(False)), int(dec.ready(True)), dec.decoded_blocks(), dec.total_blocks(),
               int(dec.payload().size))
        assert got == tuple(int(v) for v in steps[k]), (name, k)
    assert np.array_equal(dec.offsets(), g[f"{name}_offsets"])
    assert np.array_equal(dec.payload(), g[f"{name}_payload"])


@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("dims", [(48, 40, 0), (130, 66, 1), (720, 480, 0), (1024, 600, 1)])
def test_layout_matches_reference_library_directly(dims):
    """Extra sizes straight against the reference bitstream layer, where it
    was built (this container)."""
    import ctypes as C
    w, h, ch = dims
    R = O.ref()
    count8 = C.c_int()
    band = np.zeros((3, 5, 4, 4), dtype=np.int32)
    mapping = np.zeros((1 << 16, 4), dtype=np.int32)
    n32 = R.ref_layout(w, h, ch, C.byref(count8), O.ptr(band), O.ptr(mapping), 1 << 16)
    L = O.layout(w, h, ch)
    assert (L.count8, L.count32) == (count8.value, n32)
    assert L.per_subdivision == R.ref_block_count_per_subdivision(n32)
