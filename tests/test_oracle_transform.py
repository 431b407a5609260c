"""The oracle's 9/7 transform pinned by mathematics, not by its own formula:
the published CDF 9/7 analysis filters, perfect reconstruction, vanishing
moments, band orientation, and the FP32 error against FP64.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O

# CDF 9/7 analysis filters of ITU-T T.800 (JPEG 2000 irreversible transform),
# normalised to DC gain 1 (low) and Nyquist gain 2 (high), which is the
# lifting scheme with low * 1/K and high * K.
H_LOW = {0: 0.602949018236358, 1: 0.266864118442872, 2: -0.078223266528988, 3: -0.016864118442875,
         4: 0.026748757410810}
H_HIGH = {0: 1.115087052456994, 1: -0.591271763114247, 2: -0.057543526228500, 3: 0.091271763114249}


def _sym(i, n):
    """Whole-sample symmetric index."""
    while i < 0 or i >= n:
        i = -i if i < 0 else 2 * (n - 1) - i
    return i


def _analysis_1d(x):
    n = x.shape[-1]
    lo = np.zeros(x.shape[:-1] + (n // 2,))
    hi = np.zeros_like(lo)
    for k in range(n // 2):
        for t, c in H_LOW.items():
            for s in {t, -t}:
                lo[..., k] += c * x[..., _sym(2 * k + s, n)]
        for t, c in H_HIGH.items():
            for s in {t, -t}:
                hi[..., k] += c * x[..., _sym(2 * k + 1 + s, n)]
    return lo, hi


def _brute_force_2d(x):
    """Separable convolution with the textbook filters: rows (horizontal)
    first, then columns.  Bands: LL, HL (horizontal high), LH, HH."""
    lo, hi = _analysis_1d(x)
    ll, lh = (a.T for a in _analysis_1d(lo.T))
    hl, hh = (a.T for a in _analysis_1d(hi.T))
    return ll, hl, lh, hh


@pytest.mark.parametrize("n", [16, 32, 64])
def test_dwt_level_equals_textbook_filters(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n))
    got = O.dwt_level(x, n, n, n, n, f64=True)
    exp = _brute_force_2d(x)
    for g, e, name in zip(got, exp, ("LL", "HL", "LH", "HH")):
        assert np.abs(g - e).max() < 1e-12, name


@pytest.mark.parametrize("n", [16, 48, 128])
def test_perfect_reconstruction_f64(n):
    rng = np.random.default_rng(1 + n)
    x = rng.standard_normal((n, n))
    y = O.idwt_level(O.dwt_level(x, n, n, n, n, f64=True), f64=True)
    assert np.abs(y - x).max() < 1e-10


def test_constant_frame_leaves_high_bands_zero():
    L = O.layout(200, 136, 0)
    planes = [np.full(O.plane_shape(L, c), v, np.uint8) for c, v in zip(range(3), (37, 128, 201))]
    c64 = O.dwt_frame(L, planes, f64=True)
    for c in range(3):
        for l in range(5):
            for b in range(4):
                B = L.band[c][l][b]
                if B.present and b != 0:
                    assert np.abs(O.band_view(L, c64, c, l, b)).max() < 1e-12, (c, l, b)
    ll = O.band_view(L, c64, 0, 4, 0)
    # DC gain 1 per dimension and level: LL = value - 0.5 everywhere.
    assert np.abs(ll - (37 / 255.0 - 0.5)).max() < 1e-12


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_vanishing_moments(deg):
    """Polynomials up to degree 3 give zero high-pass output away from the
    edges (the 9/7 analysis high-pass has four vanishing moments)."""
    n = 64
    t = np.arange(n, dtype=np.float64) / n
    p = sum((k + 1) * t ** k for k in range(deg + 1))
    x = np.tile(p, (n, 1))  # varies along x only
    ll, hl, lh, hh = O.dwt_level(x, n, n, n, n, f64=True)
    assert np.abs(hl[:, 3:-3]).max() < 1e-12
    assert np.abs(lh).max() < 1e-12 and np.abs(hh).max() < 1e-12
    ll, hl, lh, hh = O.dwt_level(x.T.copy(), n, n, n, n, f64=True)
    assert np.abs(lh[3:-3, :]).max() < 1e-12


def test_band_orientation():
    """Horizontal high frequency lands in band 1 (HL), vertical in band 2 (LH)."""
    n = 64
    xs = np.arange(n)
    x = np.tile(np.cos(np.pi * 0.9 * xs), (n, 1))
    ll, hl, lh, hh = O.dwt_level(x, n, n, n, n, f64=True)
    e = [float((b ** 2).sum()) for b in (ll, hl, lh, hh)]
    assert e[1] > 100 * max(e[0], e[2], e[3])
    ll, hl, lh, hh = O.dwt_level(x.T.copy(), n, n, n, n, f64=True)
    e = [float((b ** 2).sum()) for b in (ll, hl, lh, hh)]
    assert e[2] > 100 * max(e[0], e[1], e[3])


def test_forward_extension_is_mirror_of_clamped_signal():
    """The encoder's extension from R valid samples to the 32-aligned domain A
    (dwt.comp:27-36): one-sided samples beyond R repeat R-1 up to A, and the
    transform is whole-sample symmetric about 0 and A-1.  Equivalent to
    padding by clamping and transforming the padded signal."""
    rng = np.random.default_rng(3)
    R, A = 40, 64
    x = rng.standard_normal((A, A))
    xr = x[:R, :R]
    padded = np.pad(xr, ((0, A - R), (0, A - R)), mode="edge")
    got = O.dwt_level(np.ascontiguousarray(xr), R, R, A, A, f64=True)
    exp = O.dwt_level(padded, A, A, A, A, f64=True)
    for g, e in zip(got, exp):
        assert np.abs(g - e).max() < 1e-12


def test_whole_frame_perfect_reconstruction_f64():
    """Five levels forward and back in FP64 reproduce the input (the inverse
    adds the 0.5 DC shift back: output in [0, 1])."""
    L = O.layout(160, 96, 1)
    rng = np.random.default_rng(5)
    planes = [rng.integers(0, 256, O.plane_shape(L, c), dtype=np.uint8) for c in range(3)]
    c64 = O.dwt_frame(L, planes, f64=True)
    rec = O.idwt_frame_f64(L, c64)
    for p, r in zip(planes, rec):
        assert np.abs(r - p / 255.0).max() < 1e-10


def test_fp32_transform_close_to_fp64():
    L = O.layout(256, 144, 0)
    rng = np.random.default_rng(9)
    planes = [rng.integers(0, 256, O.plane_shape(L, c), dtype=np.uint8) for c in range(3)]
    c32 = O.dwt_frame(L, planes, precision=2)
    c64 = O.dwt_frame(L, planes, f64=True)
    assert np.abs(c32 - c64).max() < 2e-5
    u8 = O.idwt_frame(L, c32)
    for p, r in zip(planes, u8):
        assert np.abs(p.astype(int) - r.astype(int)).max() <= 1


def _band(L, coeffs, c, l, b):
    B = L.b(c, l, b)
    return coeffs[B.coeff_off:B.coeff_off + B.bw * B.bh].reshape(B.bh, B.bw)


@pytest.mark.parametrize("w,h,chroma", [(160, 96, 0), (98, 70, 1)])
def test_fp16_ll_storage_is_levelwise_fp16_rounding(w, h, chroma):
    """precision 3 (every band of levels 0-1 stored as fp16, the LL included)
    equals an FP32 level-by-level transform whose level 0-1 outputs are
    rounded by numpy's float16 before the next level reads the LL; precision
    1 rounds only the high bands, so its level-2 input is the FP32 LL."""
    L = O.layout(w, h, chroma)
    rng = np.random.default_rng(13)
    planes = [rng.integers(0, 256, O.plane_shape(L, c), dtype=np.uint8) for c in range(3)]
    c2 = O.dwt_frame(L, planes, precision=2)
    c1 = O.dwt_frame(L, planes, precision=1)
    c3 = O.dwt_frame(L, planes, precision=3)
    f16 = lambda a: a.astype(np.float16).astype(np.float32)
    for c in range(3):
        first = 1 if (c and chroma == 0) else 0
        # The first level reads the plane: identical in every mode before storage.
        bands = [_band(L, c2, c, first, b) for b in range(4)]
        for l in range(first, O.LEVELS):
            if l > first:
                ll = bands[0]
                bands = O.dwt_level(ll, ll.shape[1], ll.shape[0], ll.shape[1], ll.shape[0], f64=False)
            if l < 2:
                bands = [f16(b) for b in bands]
            for b in range(4):
                got = _band(L, c3, c, l, b)
                assert np.array_equal(got.view(np.uint32), bands[b].view(np.uint32)), (c, l, b)
        # precision 1 keeps the LL of levels 0-1 in FP32.
        ll1 = _band(L, c1, c, first, 0)
        assert np.array_equal(ll1, _band(L, c2, c, first, 0))
    assert not np.array_equal(c1, c3)
    # The inverse rounds the LL it wrote for levels 0-1: the round trip stays within one code value.
    for p, r in zip(planes, O.idwt_frame(L, c3, precision=3)):
        assert np.abs(p.astype(int) - r.astype(int)).max() <= 1
