"""The oracle's FP16-arithmetic transform (the reference's precision 0,
shaders/dwt_common.h:29-34) pinned on the CPU: its fp16 rounding against
numpy's float16 conversion, one level against an independent numpy float16
lifting, the whole transform against the FP64 oracle within stated bounds, and
the precision-0 codec's round trip."""
import numpy as np
import pytest

from oracle import oracle as O
from pyrowave_amd import synth

ALPHA, BETA, GAMMA, DELTA = -1.586134342059924, -0.052980118572961, 0.882911075530934, 0.443506852043971
K = 1.230174104914001
K32, INV_K32 = np.float32(K), np.float32(1.0 / K)


def f16(v):
    return np.asarray(v, np.float64).astype(np.float16).astype(np.float64)


def test_f64_to_f16_rounding_matches_numpy():
    halves = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16).view(np.float16)
    finite = np.unique(halves[np.isfinite(halves)].astype(np.float64))
    mids = finite[:-1] + np.diff(finite) / 2
    rng = np.random.default_rng(2)
    xs = np.concatenate([finite, mids, np.nextafter(mids, np.inf), np.nextafter(mids, -np.inf),
                         rng.standard_normal(100000) * 4.0, rng.standard_normal(50000) * 1e-5])
    lib = O.lib()
    got = np.array([lib.or_round_f64_to_f16(float(x)) for x in xs], np.float64)
    assert np.array_equal(got, f16(xs))


def _lift_fwd_np(x, A):
    """Lifting along the last axis of x (positions -4 .. A + 3 at index p + 4):
    each step an fp16 add, then c * s + v exact and rounded once to fp16."""
    c = [float(f16(np.float32(v))) for v in (ALPHA, BETA, GAMMA, DELTA)]
    for k, start, stop in ((0, -3, A + 1), (1, -2, A), (2, -1, A - 1), (3, 0, A - 2)):
        p = np.arange(start, stop + 1, 2) + 4
        s = f16(x[..., p - 1] + x[..., p + 1])
        x[..., p] = f16(c[k] * s + x[..., p])
    return x


def _lift_inv_np(x, N):
    c = [-float(f16(np.float32(v))) for v in (ALPHA, BETA, GAMMA, DELTA)]
    for k, start, stop in ((3, -2, N + 2), (2, -1, N + 1), (1, 0, N), (0, 1, N - 1)):
        p = np.arange(start, stop + 1, 2) + 4
        s = f16(x[..., p - 1] + x[..., p + 1])
        x[..., p] = f16(c[k] * s + x[..., p])
    return x


def _scale(v, k):
    """fp16 value times the FP32 constant, the exact product rounded once."""
    return f16(np.asarray(v, np.float64) * np.float64(k))


def _dwt_level_np(img, R_x, R_y, A_x, A_y):
    ext = O.lib().or_fwd_ext
    xm = np.array([ext(p, R_x, A_x) for p in range(-4, A_x + 4)])
    ym = np.array([ext(p, R_y, A_y) for p in range(-4, A_y + 4)])
    rows = _lift_fwd_np(img[:, xm].astype(np.float64), A_x)[:, 4:4 + A_x]
    h = np.concatenate([_scale(rows[:, 0::2], INV_K32), _scale(rows[:, 1::2], K32)], axis=1)
    cols = _lift_fwd_np(h[ym, :].T.copy(), A_y)[:, 4:4 + A_y].T
    lo, hi = _scale(cols[0::2], INV_K32), _scale(cols[1::2], K32)
    W2 = A_x // 2
    return lo[:, :W2], lo[:, W2:], hi[:, :W2], hi[:, W2:]


def _isym(p, N):
    while p < 0 or p >= N:
        p = -p if p < 0 else 2 * (N - 1) - p
    return p


def _idwt_level_np(ll, hl, lh, hh):
    bh, bw = ll.shape
    W, H = 2 * bw, 2 * bh
    inter = np.zeros((H, W))
    inter[0::2, 0::2], inter[0::2, 1::2], inter[1::2, 0::2], inter[1::2, 1::2] = ll, hl, lh, hh
    ym = np.array([_isym(p, H) for p in range(-4, H + 4)])
    xm = np.array([_isym(p, W) for p in range(-4, W + 4)])
    par = (np.arange(-4, H + 4) & 1)[:, None]
    v = np.where(par == 1, _scale(inter[ym, :], INV_K32), _scale(inter[ym, :], K32))
    cols = _lift_inv_np(v.T.copy(), H)[:, 4:4 + H].T
    par = (np.arange(-4, W + 4) & 1)[None, :]
    v = np.where(par == 1, _scale(cols[:, xm], INV_K32), _scale(cols[:, xm], K32))
    return _lift_inv_np(v, W)[:, 4:4 + W]


@pytest.mark.parametrize("shape", [(48, 64, 48, 64), (37, 50, 64, 64)])
def test_fp16_level_matches_numpy_float16(shape):
    R_y, R_x, A_y, A_x = shape
    rng = np.random.default_rng(R_x)
    img = f16(rng.uniform(-0.5, 0.5, (R_y, R_x)))
    got = O.lib()
    outs = [np.zeros((A_y // 2, A_x // 2), np.float32) for _ in range(4)]
    src = np.ascontiguousarray(img, np.float32)
    got.or_dwt_level_h(O.ptr(src), R_x, R_x, R_y, A_x, A_y, *[O.ptr(o) for o in outs], A_x // 2)
    exp = _dwt_level_np(img, R_x, R_y, A_x, A_y)
    for g, e in zip(outs, exp):
        assert np.array_equal(g.astype(np.float64), e)
    # Inverse of the same bands.
    bands = [np.ascontiguousarray(b, np.float32) for b in outs]
    bh, bw = bands[0].shape
    out = np.zeros((2 * bh, 2 * bw), np.float32)
    got.or_idwt_level_h(*[O.ptr(b) for b in bands], bw, bw, bh, O.ptr(out), 2 * bw)
    assert np.array_equal(out.astype(np.float64), _idwt_level_np(*[b.astype(np.float64) for b in bands]))


# Stated bounds of the fp16 transform against the FP64 one, per band of a
# 4:2:0 scene: max-abs error per level (coefficients of the [-0.5, 0.5]
# signal) and RMS.
FP16_MAXABS = {0: 4e-3, 1: 6e-3, 2: 1.2e-2, 3: 2.5e-2, 4: 5e-2}
FP16_RMS = {0: 8e-4, 1: 1.2e-3, 2: 2.5e-3, 3: 5e-3, 4: 1e-2}


def test_fp16_transform_within_bounds_of_fp64():
    w, h = 256, 160
    L = O.layout(w, h, 0)
    planes = synth.frame(w, h, True, seed=3)
    c16 = O.dwt_frame(L, planes, precision=0).astype(np.float64)
    c64 = O.dwt_frame(L, planes, f64=True)
    assert np.array_equal(c16, f16(c16)), "every coefficient is an fp16 value"
    worst = {}
    for c in range(3):
        for l in range(5):
            for b in range(4):
                B = L.band[c][l][b]
                if not B.present:
                    continue
                d = O.band_view(L, c16, c, l, b) - O.band_view(L, c64, c, l, b)
                mx, rms = float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))
                assert mx <= FP16_MAXABS[l] and rms <= FP16_RMS[l], (c, l, b, mx, rms)
                worst[l] = max(worst.get(l, 0.0), mx)
    # fp16 math is measurably coarser than FP32 storage of the same transform.
    c32 = O.dwt_frame(L, planes, precision=2).astype(np.float64)
    assert np.abs(c16 - c64).max() > 4 * np.abs(c32 - c64).max()


@pytest.mark.parametrize("chroma", [0, 1])
def test_fp16_codec_round_trip(chroma):
    """pyrowave_c_test.cpp:220-383 at precision 0: 34x30 ramps at 64 KiB, where
    nothing is truncated, decode within +-4: the precision-0 quantizer's step
    at q = 0 is four times the precision-1 step (1/64 of the signal range for
    the level-0 HH band, about 4 code values), and the fp16 arithmetic adds
    its rounding (the reference test's +-1 is for precision 1)."""
    from helpers import ramp_planes
    L = O.layout(34, 30, chroma, precision=0)
    planes = ramp_planes(34, 30, chroma)
    bs, meta, _ = O.encode(L, planes, 64 * 1024, precision=0)
    out, pk = O.packetize(L, meta, bs, 1, 64 * 1024)
    dec = O.decode(L, [out[o:o + s].tobytes() for o, s in pk], precision=0)
    for p, r in zip(planes, dec):
        assert np.abs(p.astype(int) - r.astype(int)).max() <= 4


def test_precision0_quantizer_policy():
    """Precision 0 quantizes two bits coarser with a 512 cap
    (pyrowave_encoder.cpp:144-172), and the policy travels with the layout."""
    lib = O.lib()
    for l in range(5):
        for c in range(3):
            for b in range(4):
                r1 = lib.or_quant_resolution_policy(l, c, b, 0)
                r0 = lib.or_quant_resolution_policy(l, c, b, 1)
                assert r0 == min(512.0, r1 / 4 if r1 < 4096 else 2.0 ** (np.log2(r1) - 2)) or r0 == 512.0
    w, h = 128, 96
    planes = synth.frame(w, h, True, seed=5)
    L1, L0 = O.layout(w, h, 0), O.layout(w, h, 0, precision=0)
    coeffs = O.dwt_frame(L1, planes, precision=0)
    q1 = O.quant_analyze(L1, coeffs)
    q0 = O.quant_analyze(L0, coeffs)
    assert q0["consumed"] < q1["consumed"]
