"""The oracle's quantizer statistics, scale codes, rate-distortion buckets and
perceptual weights, pinned by what the mathematics fixes rather than by the
formulas the oracle restates.  CPU only.

Each check_* function recomputes one oracle output from an independent route:
- distortion tables (wavelet_quant.comp:113-165, 240-256) from the actual
  reconstruction error of the already-pinned pack -> parse -> dequant chain at
  every quant level;
- payload costs (wavelet_quant.comp:134-165, analyze_rate_control.comp:169-215)
  from the structural size of every packed block at every quant level
  (validate_bitstream, pyrowave_encoder.cpp:960-1046);
- quant-level truncation from the closed-form deadzone bound at step 2^q * D;
- scale codes (wavelet_quant.comp:95-111) from exact rational arithmetic: the
  smallest code whose scale keeps the block's peak under 2^e - 1/4;
- buckets (analyze_rate_control.comp:73-99) from float64 log2 and a literal
  run of the shader's Hillis-Steele max-scan;
- the distortion weight (pyrowave_encoder.cpp:107-142) from the published
  Mannos-Sakrison contrast sensitivity function at the band midpoints;
- the band term of the quant resolution (pyrowave_encoder.cpp:152-172) from
  the measured analysis gains of the FP64 CDF 9/7 transform, and the whole
  resolution (band, level and 4:2:0 chroma terms, base bits) from the L2 norms
  of the FP64 synthesis bases (noise-power normalization).

tests/test_oracle_mutants.py plants a mistake in each of these oracle
functions and requires one of these checks to fail.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from pyrowave_amd import synth

F16_ULP_REL = 2.0 ** -10


def _scene(w, h, chroma, seed):
    planes = synth.SceneGenerator(w, h, chroma == 0, seed=seed, frames=1, checker=0.1).frame(0)
    # A little extra fine noise gives blocks of every plane count.
    rng = np.random.default_rng(seed)
    return [np.clip(p.astype(np.int32) + rng.integers(-6, 7, p.shape), 0, 255).astype(np.uint8) for p in planes]


CASES = [(96, 64, 1, 5), (128, 96, 0, 9)]


def _bands(L):
    for c in range(3):
        for l in range(5):
            for b in range(4):
                B = L.band[c][l][b]
                if B.present:
                    yield c, l, b, B


def _padded(L, coeffs, c, l, b):
    B = L.band[c][l][b]
    out = np.zeros((B.by8 * 8, B.bx8 * 8), np.float64)
    out[:B.bh, :B.bw] = O.band_view(L, coeffs, c, l, b)
    return out


def _blocks8(a):
    """(by8, bx8, 64) view of 8x8 blocks of a padded band."""
    h, w = a.shape
    return a.reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3).reshape(h // 8, w // 8, 64)


def _decode_at_quant(L, coeffs, q):
    quant = np.full(L.count32, q, np.int32)
    bs, meta = O.pack(L, coeffs, quant, 1)
    out, pk = O.packetize(L, meta, bs, 1, 1 << 30)
    d = O.Decoder(L)
    assert d.push(out[:int(pk[0][1])].tobytes())
    return O.dequant(L, d.payload(), d.offsets()), bs, meta


def _stats(L, coeffs):
    """Per band: arrays cost[by8, bx8, 16], dist[..., 16], code word, msb."""
    res = {}
    for c, l, b, B in _bands(L):
        cost = np.zeros((B.by8, B.bx8, 16), np.int64)
        dist = np.zeros((B.by8, B.bx8, 16), np.float64)
        cw = np.zeros((B.by8, B.bx8), np.int64)
        msb = np.zeros((B.by8, B.bx8), np.int64)
        for y in range(B.by8):
            for x in range(B.bx8):
                k, d, w, m = O.block8_stats(L, coeffs, c, l, b, x, y)
                cost[y, x], dist[y, x], cw[y, x], msb[y, x] = k, d, w, m
        res[(c, l, b)] = (cost, dist, cw, msb)
    return res


def _frame(case):
    w, h, ch, seed = case
    L = O.layout(w, h, ch)
    coeffs = O.dwt_frame(L, _scene(w, h, ch, seed))
    return L, coeffs


def C_L(L):
    return ctypes.byref(L)


def _f16(x):
    return np.minimum(x, 60000.0).astype(np.float16).astype(np.float64)


# ---------------------------------------------------------------------------
# checks (each raises AssertionError on a mismatch)
# ---------------------------------------------------------------------------

def check_distortion_through_dequant(case=CASES[0]):
    """errors[q].square_error for q >= 1 is the fp16-stored squared error of the
    block reconstructed at quant level q, in coefficient units, times the
    band's distortion weight; past the block's last plane (the 'discard'
    entry, wavelet_quant.comp:253-256) the reference omits the weight.
    Reconstructions come from pack -> packetize -> parse -> dequant."""
    L, coeffs = _frame(case)
    st = _stats(L, coeffs)
    checked = 0
    for q in range(1, 16):
        rec, _, _ = _decode_at_quant(L, coeffs, q)
        for c, l, b, B in _bands(L):
            cost, dist, cw, msb = st[(c, l, b)]
            err = _blocks8(_padded(L, coeffs, c, l, b) - _padded(L, rec, c, l, b))
            se = (err * err).sum(axis=2)
            rdo = float(O.lib().or_rdo_distortion_scale(C_L(L), l, c, b))
            weighted = np.where(q <= msb, se * rdo, se)
            exp = np.where(msb < 0, 0.0, _f16(weighted))
            got = dist[:, :, q]
            tol = 2 * F16_ULP_REL * np.abs(exp) + 2.0 ** -23
            bad = np.abs(got - exp) > tol
            assert not bad.any(), (q, (c, l, b), got[bad][:4], exp[bad][:4])
            checked += int((msb >= 0).sum())
    assert checked > 1000
    return True


def check_dist32_is_sum_of_blocks(case=CASES[1]):
    """The 32x32 distortion per quant level is the sum of its 16 8x8 entries
    (analyze_rate_control.comp:169-215; out-of-band 8x8 blocks add 0)."""
    L, coeffs = _frame(case)
    qa = O.quant_analyze(L, coeffs, with_dist=True)
    st = _stats(L, coeffs)
    for c, l, b, B in _bands(L):
        dist = st[(c, l, b)][1]
        pad = np.zeros((B.by32 * 4, B.bx32 * 4, 16))
        pad[:B.by8, :B.bx8] = dist
        s = pad.reshape(B.by32, 4, B.bx32, 4, 16).sum(axis=(1, 3)).reshape(-1, 16)
        got = qa["dist32"][B.off32:B.off32 + B.bx32 * B.by32].astype(np.float64)
        assert np.allclose(got, s, rtol=1e-6, atol=1e-6), (c, l, b)
    assert (qa["dist32"][:, 0] == 0).all()
    return True


def check_cost_through_pack(case=CASES[0]):
    """At every quant level the analyzed words of each 32x32 block equal the
    words of the block as packed at that level, and every packed block's
    structure (header, control words, planes, signs) accounts for exactly that
    many words."""
    L, coeffs = _frame(case)
    qa = O.quant_analyze(L, coeffs)
    for q in range(16):
        quant = np.full(L.count32, q, np.int32)
        bs, meta = O.pack(L, coeffs, quant, 1)
        assert np.array_equal(meta[:, 1], qa["words"][:, q].astype(np.uint32)), q
        for blk in range(L.count32):
            off, n = int(meta[blk, 0]), int(meta[blk, 1])
            if n:
                assert O.validate_block(L, bs[off:off + n], blk, n), (q, blk)
    return True


def check_truncation_bound(case=CASES[1]):
    """Deadzone reconstruction at quant level q: with D the block's step
    (decode_quant(code) * (scale / 8 + 1/4), bitstream.md:437-470), a zero
    means |c| < 2^q D and otherwise |c - c'| <= 2^q D / 2 with the sign kept."""
    L, coeffs = _frame(case)
    st = _stats(L, coeffs)
    lib = O.lib()
    for q in (0, 1, 2, 3, 5):
        rec, _, _ = _decode_at_quant(L, coeffs, q)
        for c, l, b, B in _bands(L):
            cost, dist, cw, msb = st[(c, l, b)]
            band_code = int(lib.or_encode_quant(np.float32(1.0) / np.float32(lib.or_quant_resolution(l, c, b))))
            step0 = float(lib.or_decode_quant(band_code))
            scale = ((cw >> 20) & 15).astype(np.float64) / 8.0 + 0.25
            step = (step0 * scale * 2.0 ** q)[:, :, None]
            src = _blocks8(_padded(L, coeffs, c, l, b))
            got = _blocks8(_padded(L, rec, c, l, b))
            live = (msb >= 0)[:, :, None] & np.ones((1, 1, 64), bool)
            z = (got == 0) & live
            nz = (got != 0) & live
            stepb = np.broadcast_to(step, src.shape)
            assert (np.abs(src[z]) < stepb[z] * (1 + 1e-6)).all(), (q, c, l, b)
            assert (np.abs(src[nz] - got[nz]) <= stepb[nz] / 2 * (1 + 1e-5) + 1e-7).all(), (q, c, l, b)
            assert (np.sign(src[nz]) == np.sign(got[nz])).all()
    return True


def check_scale_codes(case=CASES[0]):
    """Scale code = the smallest c with M / (c / 8 + 1/4) <= 2^e - 1/4, where M
    is the block's peak |coefficient * resolution| and 2^(e-1) <= M - 1/4 < 2^e
    (exact rationals; ties within float rounding are skipped)."""
    L, coeffs = _frame(case)
    st = _stats(L, coeffs)
    lib = O.lib()
    checked = 0
    for c, l, b, B in _bands(L):
        cost, dist, cw, msb = st[(c, l, b)]
        qres = np.float32(1.0) / np.float32(lib.or_decode_quant(lib.or_encode_quant(
            np.float32(1.0) / np.float32(lib.or_quant_resolution(l, c, b)))))
        band = _padded(L, coeffs, c, l, b).astype(np.float32)
        t = np.abs(band * qres)
        M = _blocks8(t.astype(np.float64)).max(axis=2)
        for y in range(B.by8):
            for x in range(B.bx8):
                m = Fraction(float(np.float32(min(M[y, x], 16383.75))))
                if m < 1:
                    assert cw[y, x] == 0
                    continue
                e = math.floor(math.log2(m - Fraction(1, 4))) + 1
                while Fraction(2) ** (e - 1) > m - Fraction(1, 4):
                    e -= 1
                while Fraction(2) ** e <= m - Fraction(1, 4):
                    e += 1
                target = Fraction(2) ** e - Fraction(1, 4)
                x8 = 8 * (m / target - Fraction(1, 4))
                if abs(x8 - round(x8)) < Fraction(1, 10 ** 5):
                    continue
                code = math.ceil(x8)
                assert 0 <= code <= 15
                peak = m / (Fraction(code, 8) + Fraction(1, 4))
                if abs(peak - 1) < Fraction(1, 10 ** 5):
                    continue
                if peak < 1:
                    # The scaled block quantizes to zero: no code word.
                    assert cw[y, x] == 0
                    continue
                assert int((cw[y, x] >> 20) & 15) == code, ((c, l, b), x, y, float(m), code)
                checked += 1
    assert checked > 300
    return True


def _bucket_f64(r):
    if r <= 0:
        return 0
    v = 60.5 + 2.0 * math.log2(r)
    return max(math.floor(v), 0)


def _near_tie(r):
    if r <= 0:
        return False
    v = 60.5 + 2.0 * math.log2(r)
    return abs(v - round(v)) < 1e-5


def check_bucket_of_ratio():
    """uint(max(60 + 2 log2(r) + 0.5, 0)) against float64 log2, away from ties."""
    rng = np.random.default_rng(3)
    rs = np.float32(2.0) ** rng.uniform(-45, 45, 20000).astype(np.float32)
    for r in list(rs) + [np.float32(1.0), np.float32(0.5), np.float32(1e-30), np.float32(3e4)]:
        if _near_tie(float(r)):
            continue
        assert O.bucket_of_ratio(r) == _bucket_f64(float(r)), float(r)
    assert O.bucket_of_ratio(0.0) == 0
    return True


def _maxscan_literal(v):
    """inclusive_max_clustered16 (analyze_rate_control.comp:85-99), run lane by
    lane: clamp to 128 - 16 + lane, then shuffle-up steps 1, 2, 4, 8 adding
    the distance."""
    v = [min(x, 128 - 16 + lane) for lane, x in enumerate(v)]
    i = 1
    while i < 16:
        up = [v[lane - i] + i if lane >= i else 0 for lane in range(16)]
        v = [max(v[lane], up[lane]) for lane in range(16)]
        i *= 2
    return v


def check_bucket_scan():
    """The oracle's bucket scan against the literal shader loop on random raw
    buckets over the whole u32-reachable range, including the clamp."""
    rng = np.random.default_rng(11)
    for i in range(3000):
        raw = rng.integers(0, 200 if i % 2 else 130, 16)
        raw[0] = 0
        assert O.bucket_scan(raw) == _maxscan_literal([int(x) for x in raw]), raw
    return True


def check_bucket_tables(case=CASES[1]):
    """Per 32x32 block: buckets from float64 log2 of (D_q - D_0) / (W_0 - W_q)
    (0 where the words do not change), then the literal max-scan."""
    L, coeffs = _frame(case)
    qa = O.quant_analyze(L, coeffs, with_dist=True)
    W = qa["words"].astype(np.int64)
    D = qa["dist32"].astype(np.float64)
    checked = 0
    for blk in range(L.count32):
        raw = [0] * 16
        tie = False
        for q in range(1, 16):
            if W[blk, q] == W[blk, 0]:
                continue
            r = float(np.float32(np.float32(max(D[blk, q] - D[blk, 0], 0.0)) / np.float32(W[blk, 0] - W[blk, q])))
            tie |= _near_tie(r)
            raw[q] = _bucket_f64(r)
        if tie:
            continue
        assert list(qa["bucket"][blk].astype(int)) == _maxscan_literal(raw), blk
        checked += 1
    assert checked > 0.9 * L.count32
    return True


def check_totals_and_consumed(case=CASES[1]):
    """Savings per (bucket, subdivision) and the consumed words, recomputed
    from the word and bucket tables (analyze_rate_control.comp:127-146)."""
    L, coeffs = _frame(case)
    qa = O.quant_analyze(L, coeffs)
    W = qa["words"].astype(np.int64)
    shamt = int(L.per_subdivision).bit_length() - 1
    totals = np.zeros(2048, np.int64)
    for blk in range(L.count32):
        for q in range(1, 16):
            s = W[blk, q - 1] - W[blk, q]
            if s:
                totals[int(qa["bucket"][blk, q]) * 16 + (blk >> shamt)] += s
    assert np.array_equal(totals, qa["totals"].astype(np.int64))
    assert qa["consumed"] == int(W[:, 0].sum())
    return True


def mannos_sakrison(f):
    """Mannos & Sakrison (1974) contrast sensitivity, f in cycles/degree."""
    return 2.6 * (0.0192 + 0.114 * f) * math.exp(-(0.114 * f) ** 1.1)


def _analysis_band_gains():
    """L1 norm of each level-0 analysis basis (peak output for a +-1 input),
    from impulse responses of the FP64 transform."""
    N = 32
    g = np.zeros(4)
    for y in range(N):
        for x in range(N):
            img = np.zeros((N, N))
            img[y, x] = 1.0
            bands = O.dwt_level(img, N, N, N, N)
            for b in range(4):
                g[b] += abs(bands[b][N // 4, N // 4])
    return g


def check_quant_resolution_band_term():
    """The resolution's band term keeps the LL, HL/LH and HH integer ranges
    aligned: resolution x analysis gain is the same power of two for the four
    bands of a level (log2 gains 0.93 / 1.84 / 2.75 against +2 / +1 / 0 bits),
    at every level and component before the cap."""
    g = _analysis_band_gains()
    lib = O.lib()
    for l in range(4):
        for c in range(3):
            bits = [math.log2(lib.or_quant_resolution(l, c, b)) for b in range(4)]
            if max(bits) >= 12:
                continue
            ref = bits[3] + math.log2(g[3])
            for b in range(3):
                assert abs(bits[b] + math.log2(g[b]) - ref) < 0.35, (l, c, b, bits, g)
    return True


def _synthesis_norms(N=256):
    """log2 of the L2 norm of the synthesis basis of band b after s analysis
    stages (s = 1..5): a unit impulse at the band's centre, run through one
    FP64 inverse level and then s - 1 inverse levels of LL only."""
    out = {}
    for s in range(1, 6):
        for b in range(4):
            n = N >> s
            bands = [np.zeros((n, n)) for _ in range(4)]
            bands[b][n // 2, n // 2] = 1.0
            x = O.idwt_level(bands)
            while x.shape[0] < N:
                m = x.shape[0]
                z = np.zeros((m, m))
                x = O.idwt_level([x, z, z, z])
            out[s, b] = 0.5 * math.log2(float((x ** 2).sum()))
    return out


def check_quant_resolution_noise_power():
    """The quant resolution is noise-power normalized (pyrowave_encoder.cpp:
    152-172): a quantization step of 1/resolution in any band reconstructs with
    the same L2 norm, 2^-9 (half of one 8-bit step) for the precision >= 1
    policy and 2^-7 for precision 0.  So resolution = 2^(9 + log2 of the
    synthesis basis norm), rounded to a power of two, with the norm measured
    on the FP64 transform at the band's stage count in its own plane (read
    from the layout: a 4:2:0 chroma band at level l has passed l stages, a
    luma band l + 1).  This fixes the band term, the bit per level, the 4:2:0
    chroma bit and the base bits at once.  Transcribed, not fixed by this
    rule: one bit less for 4:4:4 chroma, and the caps 4096 / 512."""
    norms = _synthesis_norms()
    lib = O.lib()
    for chroma in (0, 1):
        L = O.layout(512, 512, chroma)
        for c, l, b, B in _bands(L):
            ph, pw = O.plane_shape(L, c)
            s = int(round(math.log2(pw / B.bw)))
            assert pw == B.bw << s
            for precision0, base, cap in ((0, 9, 4096.0), (1, 7, 512.0)):
                want = base + norms[s, b] - (1 if c and chroma == 1 else 0)
                got = lib.or_quant_resolution_policy(l, c, b, precision0)
                if got == cap:
                    assert want > math.log2(cap) - 0.25, (chroma, c, l, b, precision0, got, want)
                else:
                    assert got < cap and abs(math.log2(got) - want) < 0.25, (chroma, c, l, b, precision0, got, want)
    # The distortion weight takes the resolution before the cap: for a band
    # at the cap, rdo / CSF^2 * 256 is still 4^(9 + norm).
    L = O.layout(512, 512, 0)
    B = L.band[0][4][0]
    assert B.present
    csf = mannos_sakrison(max(math.hypot(0.25, 0.25) * 0.34 * 96.0 / 16.0, 8.0))
    unc = math.sqrt(float(lib.or_rdo_distortion_scale(C_L(L), 4, 0, 0)) * 256.0) / csf
    assert abs(math.log2(unc) - 9 - norms[5, 0]) < 0.25
    return True


def check_rdo_weights():
    """Distortion weight = (CSF(f) * resolution)^2 / 256 with the published
    Mannos-Sakrison CSF at the band's midpoint frequency (0.25 or 0.75 of
    Nyquist per axis, Nyquist = 0.34 * 96 cycles/degree at level 0, halving
    per level, floored at 8 cpd), chroma of subsampled frames weighted 0.6
    below the top level, and resolution = 2^(8 + band term + level - chroma)
    without the cap (pyrowave_encoder.cpp:107-142)."""
    for chroma in (0, 1):
        L = O.layout(256, 256, chroma)
        for c, l, b, B in _bands(L):
            h = 0.75 if b & 1 else 0.25
            v = 0.75 if b & 2 else 0.25
            f = max(math.hypot(h, v) * 0.34 * 96.0 * 2.0 ** -l, 8.0)
            csf = mannos_sakrison(f)
            if c and l != 4 and chroma == 0:
                csf *= 0.6
            bits = 8 + (2 if b == 0 else 1 if b < 3 else 0) + l - (1 if c else 0)
            exp = (csf * 2.0 ** bits) ** 2 / 256.0
            got = float(O.lib().or_rdo_distortion_scale(C_L(L), l, c, b))
            assert abs(got - exp) <= 2e-5 * exp, (chroma, c, l, b, got, exp)
    # The published curve peaks near 8 cycles/degree at about 1.0.
    fs = np.linspace(1, 30, 2901)
    vals = [mannos_sakrison(x) for x in fs]
    assert abs(fs[int(np.argmax(vals))] - 8.0) < 0.5 and abs(max(vals) - 1.0) < 0.02
    return True


CHECKS = [check_distortion_through_dequant, check_dist32_is_sum_of_blocks, check_cost_through_pack,
          check_truncation_bound, check_scale_codes, check_bucket_of_ratio, check_bucket_scan, check_bucket_tables,
          check_totals_and_consumed, check_quant_resolution_band_term, check_quant_resolution_noise_power,
          check_rdo_weights]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__ for c in CHECKS])
def test_oracle_pinned(check):
    assert check()


@pytest.mark.parametrize("case", CASES[1:])
def test_distortion_and_costs_420(case):
    """The distortion and cost pins on a 4:2:0 frame as well (chroma weights)."""
    assert check_distortion_through_dequant(case)
    assert check_cost_through_pack(case)
    assert check_scale_codes(case)
