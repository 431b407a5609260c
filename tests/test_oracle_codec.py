"""The oracle's integer stages pinned by closed forms, literal simulations of
the reference's GPU loops, structural validation and the reference's own
C-test tolerances.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pyrowave_amd import synth


def ramp_planes(w, h, chroma):
    """pyrowave_c_test.cpp:288-308: luma 3x+5y, Cb 7x+3y, Cr 3x+5y (mod 256)."""
    ys, xs = np.mgrid[0:h, 0:w]
    y = ((3 * xs + 5 * ys) & 255).astype(np.uint8)
    ch, cw = (h // 2, w // 2) if chroma == 0 else (h, w)
    ys, xs = np.mgrid[0:ch, 0:cw]
    return [y, ((7 * xs + 3 * ys) & 255).astype(np.uint8), ((3 * xs + 5 * ys) & 255).astype(np.uint8)]


def scene(w, h, chroma, seed=1):
    return synth.SceneGenerator(w, h, chroma == 0, seed=seed, frames=1).frame(0)


# --- fp16 -------------------------------------------------------------------

def test_f16_conversion_matches_numpy():
    lib = O.lib()
    lib.or_round_to_f16_array.argtypes = [C.c_void_p, C.c_long]
    halves = np.arange(0, 1 << 16, dtype=np.uint32).astype(np.uint16).view(np.float16)
    finite = np.unique(halves[np.isfinite(halves)].astype(np.float64))
    mids = finite[:-1] + np.diff(finite) / 2  # ties: round to even
    rng = np.random.default_rng(0)
    samples = np.concatenate([
        finite.astype(np.float32), mids.astype(np.float32),
        np.nextafter(mids.astype(np.float32), np.float32(np.inf)),
        np.nextafter(mids.astype(np.float32), np.float32(-np.inf)),
        rng.standard_normal(200000).astype(np.float32) * np.float32(3000.0),
        (rng.random(200000) * 1e-4).astype(np.float32),
    ])
    v = samples.copy()
    lib.or_round_to_f16_array(O.ptr(v), v.size)
    exp = samples.astype(np.float16).astype(np.float32)
    ok = np.isfinite(exp)
    assert np.array_equal(v[ok].view(np.uint32), exp[ok].view(np.uint32))


def test_fast_f16_rounding_exhaustive():
    """The fast rounding used for stored distortions equals the general
    conversion for every float in [0, 60000] (about 1.2e9 values)."""
    assert O.lib().or_check_round_f16_fast() == 0


# --- quant codes (pyrowave_common.hpp:124-145) --------------------------------

def test_quant_code_closed_form_and_round_trip():
    lib = O.lib()
    for code in range(256):
        e = min(code >> 3, 24)
        m = code & 7
        assert lib.or_decode_quant(code) == np.float32(2.0 ** (4 - e) * (1 + m / 8.0))
        if (code >> 3) <= 24:
            assert lib.or_encode_quant(lib.or_decode_quant(code)) == code


# --- finalize: literal simulation of analyze_rate_control_finalize.comp ---------

def finalize_literal(totals):
    """512 invocations, each holding a uvec4 of 4 consecutive totals
    (analyze_rate_control_finalize.comp:14-38): local inclusive scan of the
    uvec4, then `for (step = 1; step < 512 / 2; step *= 2)` rounds of
    read-neighbour / barrier / add / publish.  uint32 wrap-around."""
    v = totals.astype(np.uint64).reshape(512, 4).copy()
    v[:, 1] += v[:, 0]
    v[:, 2] += v[:, 1]
    v[:, 3] += v[:, 2]
    v &= 0xffffffff
    shared = v[:, 3].copy()
    step = 1
    while step < 512 // 2:
        up = np.zeros(512, dtype=np.uint64)
        up[step:] = shared[:-step]
        v = (v + up[:, None]) & 0xffffffff
        shared = v[:, 3].copy()
        step *= 2
    return v.reshape(-1).astype(np.uint32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_finalize_q1_matches_literal_loop(seed):
    rng = np.random.default_rng(seed)
    totals = rng.integers(0, 5000, 2048).astype(np.uint32)
    totals[rng.random(2048) < 0.5] = 0
    got = O.finalize(totals, q1=True)
    assert np.array_equal(got, finalize_literal(totals))
    # Below flat index 1024 the quirk is invisible: a true prefix sum there.
    assert np.array_equal(got[:1024], np.cumsum(totals.astype(np.uint64))[:1024].astype(np.uint32))


def test_finalize_exact_mode_is_prefix_sum():
    totals = np.random.default_rng(7).integers(0, 1 << 20, 2048).astype(np.uint32)
    assert np.array_equal(O.finalize(totals, q1=False), np.cumsum(totals.astype(np.uint64)).astype(np.uint32))


# --- resolve: literal simulation of resolve_rate_control.comp -----------------

def resolve_literal(L, qa, prefix, budget, subgroup=64):
    """analyze_rate_control.comp:127-146 places op {q, block | saving << 16}
    at block + bucket * num_blocks_aligned; resolve_rate_control.comp:38-79
    runs one workgroup per flat (bucket, subdivision) index and walks its
    ops in subgroup-sized chunks."""
    per_sub = L.per_subdivision
    aligned = per_sub * 16
    op_q = np.zeros(128 * aligned, dtype=np.int64)
    op_save = np.zeros(128 * aligned, dtype=np.int64)
    op_blk = np.zeros(128 * aligned, dtype=np.int64)
    words = qa["words"].astype(np.int64)
    for b in range(L.count32):
        for q in range(1, 16):
            saving = words[b, q - 1] - words[b, q]
            if saving:
                k = b + int(qa["bucket"][b, q]) * aligned
                op_q[k], op_save[k], op_blk[k] = q, saving, b
    target = budget & ~3
    target = (target - 8) // 4 if target >= 8 else 0
    quant = np.zeros(L.count32, dtype=np.int64)
    for x in range(2048):
        required = int(qa["consumed"]) - target
        if x:
            if prefix[x] == prefix[x - 1]:
                continue
            required -= int(prefix[x - 1])
        elif prefix[0] == 0:
            continue
        if required <= 0:
            continue
        total = 0
        i = 0
        while i < per_sub and total < required:
            sl = slice(x * per_sub + i, x * per_sub + min(i + subgroup, per_sub))
            save = op_save[sl]
            scan = np.cumsum(save)
            apply = (total + scan - save < required) & (save != 0)
            for q, blk in zip(op_q[sl][apply], op_blk[sl][apply]):
                quant[blk] = max(quant[blk], q)
            total += int(scan[-1])
            i += subgroup
    return quant


@pytest.mark.parametrize("dims,budgets", [((256, 144, 0), (3000, 9000, 30000)),
                                          ((160, 160, 1), (2000, 12000, 40000))])
@pytest.mark.parametrize("q1", [True, False])
def test_resolve_matches_literal_loop(dims, budgets, q1):
    w, h, ch = dims
    L = O.layout(w, h, ch)
    qa = O.quant_analyze(L, O.dwt_frame(L, scene(w, h, ch)))
    prefix = O.finalize(qa["totals"], q1=q1)
    for budget in budgets:
        got = O.resolve(L, qa, prefix, budget)
        assert np.array_equal(got, resolve_literal(L, qa, prefix, budget)), budget
        assert np.array_equal(got, resolve_literal(L, qa, prefix, budget, subgroup=32)), budget


def test_buckets_strictly_increase_over_q():
    """inclusive_max_clustered16 (analyze_rate_control.comp:85-99) leaves
    buckets strictly increasing in q and below 128 - 16 + q + 1."""
    L = O.layout(256, 144, 0)
    qa = O.quant_analyze(L, O.dwt_frame(L, scene(256, 144, 0)))
    bk = qa["bucket"].astype(int)
    assert (bk[:, 0] == 0).all()
    assert (np.diff(bk, axis=1) >= 1).all()
    assert (bk <= 112 + np.arange(16)).all()
    w = qa["words"].astype(int)
    assert (np.diff(w, axis=1) <= 0).all(), "words never grow with q"
    assert qa["consumed"] == int(w[:, 0].sum())


# --- packing, budget and round trips ------------------------------------------

def _blocks(bs, meta):
    for b, (off, n) in enumerate(meta):
        if n:
            yield b, bs[off:off + n]


@pytest.mark.parametrize("dims", [(256, 144, 0), (96, 200, 1), (34, 30, 0)])
@pytest.mark.parametrize("budget", [2500, 20000, 1 << 20])
def test_packed_blocks_validate_and_match_analyzed_words(dims, budget):
    w, h, ch = dims
    L = O.layout(w, h, ch)
    planes = scene(w, h, ch, seed=3)
    coeffs = O.dwt_frame(L, planes)
    qa = O.quant_analyze(L, coeffs)
    prefix = O.finalize(qa["totals"])
    quant = O.resolve(L, qa, prefix, budget)
    bs, meta = O.pack(L, coeffs, quant, 1)
    for b in range(L.count32):
        assert meta[b, 1] == qa["words"][b, quant[b]], b
    for b, words in _blocks(bs, meta):
        assert O.validate_block(L, words, b, int(meta[b, 1])), b
    total = 8 + 4 * int(meta[:, 1].sum())
    assert total <= max(budget, 8 + 4 * int(qa["words"][:, 15].sum()))


@pytest.mark.parametrize("budget", [4000, 8000, 16000, 32000, 64000])
def test_budget_fill(budget):
    """Bytes never exceed the budget and fill at least 95 % of it when
    truncation is needed (pyrowave_c_test.cpp:492-494)."""
    L = O.layout(320, 192, 0)
    planes = scene(320, 192, 0, seed=4)
    bs, meta, quant = O.encode(L, planes, budget)
    out, pk = O.packetize(L, meta, bs, 1, 1 << 30)
    size = int(sum(s for _, s in pk))
    unconstrained = 8 + 4 * int(O.quant_analyze(L, O.dwt_frame(L, planes))["words"][:, 0].sum())
    assert size <= budget
    if unconstrained > budget:
        assert size >= 0.95 * budget, (size, budget)


@pytest.mark.parametrize("chroma", [0, 1])
def test_reference_c_test_roundtrip(chroma):
    """pyrowave_c_test.cpp:220-383: 34x30 ramps at 64 KiB, one packet at
    offset 0, |decoded - input| <= 1 on every plane, ready/clear/re-push."""
    L = O.layout(34, 30, chroma)
    planes = ramp_planes(34, 30, chroma)
    bs, meta, _ = O.encode(L, planes, 64 * 1024)
    out, pk = O.packetize(L, meta, bs, 1, 64 * 1024)
    assert len(pk) == 1 and pk[0][0] == 0
    d = O.Decoder(L)
    data = out[:pk[0][1]].tobytes()
    assert d.push(data) and d.ready()
    d.clear()
    assert not d.ready()
    assert d.push(data) and d.ready()
    dec = O.idwt_frame(L, O.dequant(L, d.payload(), d.offsets()))
    for p, r in zip(planes, dec):
        assert np.abs(p.astype(int) - r.astype(int)).max() <= 1


def test_empty_decode_is_mid_gray():
    """pyrowave_c_test.cpp:105-162: no packets -> not ready, and decoding
    all-missing blocks gives 0x7f or 0x80."""
    L = O.layout(128, 128, 0)
    d = O.Decoder(L)
    assert not d.ready() and not d.ready(True)
    offsets = np.full(L.count32, 0xffffffff, np.uint32)
    dec = O.idwt_frame(L, O.dequant(L, np.zeros(4, np.uint32), offsets))
    for p in dec:
        assert set(np.unique(p)) <= {0x7f, 0x80}


def test_dequant_error_within_half_step():
    """Closed-form reconstruction bound at q = 0: with step
    D = decode_quant(q_code) * (scale / 8 + 0.25) of each 8x8 block, a
    decoded zero means |c| < D and otherwise |c - c'| <= D / 2 and the sign
    matches (wavelet_quant.comp deadzone quantizer, wavelet_dequant.comp:146-159,
    312-315)."""
    w, h, ch = 160, 96, 1
    L = O.layout(w, h, ch)
    coeffs = O.dwt_frame(L, scene(w, h, ch, seed=6))
    quant = np.zeros(L.count32, np.int32)
    bs, meta = O.pack(L, coeffs, quant, 1)
    d = O.Decoder(L)
    out, pk = O.packetize(L, meta, bs, 1, 1 << 30)
    assert d.push(out[:int(pk[0][1])].tobytes())
    dec = O.dequant(L, d.payload(), d.offsets())
    descs = O.block_descs(L)
    lib = O.lib()
    checked = 0
    for b, words in _blocks(bs, meta):
        raw = words.view(np.uint8)
        ballot = int(words[0]) & 0xffff
        qcode = int(raw[4])
        n = bin(ballot).count("1")
        c, l, bd = int(descs["component"][b]), int(descs["level"][b]), int(descs["band"][b])
        src = O.band_view(L, coeffs, c, l, bd)
        rec = O.band_view(L, dec, c, l, bd)
        j = 0
        for k in range(16):
            x0 = int(descs["bx"][b]) * 32 + 8 * (k & 3)
            y0 = int(descs["by"][b]) * 32 + 8 * (k >> 2)
            if x0 >= src.shape[1] or y0 >= src.shape[0]:
                continue
            s = src[y0:y0 + 8, x0:x0 + 8].astype(np.float64)
            r = rec[y0:y0 + 8, x0:x0 + 8].astype(np.float64)
            if not (ballot >> k) & 1:
                assert (r == 0).all()
                continue
            scale = int(raw[8 + 2 * n + j]) >> 4
            j += 1
            step = float(lib.or_decode_quant(qcode)) * (scale / 8.0 + 0.25)
            z = r == 0
            assert (np.abs(s[z]) < step * (1 + 1e-6)).all()
            assert (np.abs(s[~z] - r[~z]) <= step / 2 * (1 + 1e-5) + 1e-7).all()
            assert (np.sign(s[~z]) == np.sign(r[~z])).all()
            checked += int((~z).sum())
    assert checked > 1000
