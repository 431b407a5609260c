"""Pins the oracle's PSNR-HVS-M-H (oracle/psnr_hvs_m.c) on the CPU.

- The 8x8 butterfly DCT (shaders/psnr_hvs_m.comp:58-97) equals the
  orthonormal DCT-II (scipy), signs included; Parseval holds.
- Masking effect: zero for a flat block, unchanged by a DC offset, linear in
  the block's contrast.
- Tables (psnr.cpp:156-204): the contrast sensitivity function is the
  Mannos-Sakrison CSF, whose peak is about 1.0 near 8 cycles/degree; tables
  are symmetric in x and y, the mask peaks at 1 and excludes DC, the inverse
  mask is its reciprocal.
- Totals: an independent float64 implementation of the published algorithm
  (psnrhvsm.m, Ponomarenko et al.; scipy DCT, numpy sums) agrees on the
  summed errors within float32 rounding; identical images give zero; the metric is
  symmetric in its two images and grows with the distortion.
"""
import numpy as np
import pytest
import scipy.fft as F

from oracle import oracle as O


def test_dct_is_orthonormal_dct2():
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = rng.random((8, 8)).astype(np.float32)
        d = O.dct8x8(x)
        r = F.dctn(x.astype(np.float64), norm="ortho")
        assert np.abs(d - r).max() < 1e-6
        assert abs(float((d.astype(np.float64) ** 2).sum()) - float((x.astype(np.float64) ** 2).sum())) < 1e-5


def test_maskeff_properties():
    rng = np.random.default_rng(2)
    mask = O.psnr_hvs_m_tables(1080, 1.0)[1]
    flat = np.full((8, 8), 0.37, np.float32)
    assert O.maskeff(flat, O.dct8x8(flat), mask) == 0.0
    x = (rng.random((8, 8)) * 0.5).astype(np.float32)
    m = O.maskeff(x, O.dct8x8(x), mask)
    assert m > 0
    y = x + np.float32(0.25)
    assert abs(O.maskeff(y, O.dct8x8(y), mask) - m) <= 1e-5 * m
    z = x * np.float32(0.5)
    assert abs(O.maskeff(z, O.dct8x8(z), mask) - 0.5 * m) <= 1e-5 * m


def test_csf_tables():
    # Entry [0][0] is 2.6 * CSF(sqrt(2) / 16 * nyquist); nyquist grows with the height.
    f, v = [], []
    for h in range(4000, 20000, 25):
        csf = O.psnr_hvs_m_tables(h, 1.0)[0]
        f.append(np.sqrt(2) / 16 * h * np.pi / 360)
        v.append(float(csf[0, 0]) / 2.6)
    k = int(np.argmax(v))
    assert 7.0 < f[k] < 9.0, f[k]
    assert abs(v[k] - 1.0) < 0.02, v[k]
    for h in (480, 1080, 2160):
        for i in (0, 7, 15):
            csf, mask, inv = O.psnr_hvs_m_tables(h, O.psnr_height_factor(i))
            assert np.array_equal(csf, csf.T)
            assert mask[0, 0] == 0 and inv[0, 0] == 0
            assert abs(float(mask.max()) - 1.0) < 1e-6
            m = mask.reshape(-1)[1:].astype(np.float64)
            assert np.allclose(inv.reshape(-1)[1:] * m, 1.0, rtol=1e-6)
    assert [O.psnr_height_factor(i) for i in (0, 8, 15)] == [1.0, 2.0, 2.875]


def _reference_f64(a, b, nfactors):
    """PSNR-HVS-M-H from the published algorithm in float64 (8x8 DCT blocks at
    a stride of 4, zero outside the image), with the oracle's tables."""
    h, w = a.shape
    nbx, nby = (w - 1) // 4 + 1, (h - 1) // 4 + 1
    pad = lambda p: np.pad(p.astype(np.float64) / 255.0, ((0, 8), (0, 8)))
    A, B = pad(a), pad(b)
    tables = [O.psnr_hvs_m_tables(h, O.psnr_height_factor(f)) for f in range(nfactors)]

    def vari(z):
        # psnrhvsm.m: var(z(:)) * numel(z)
        return z.var(ddof=1) * z.size

    def mask_of(z, d, mc):
        m = float((d ** 2 * mc).sum())
        pop = vari(z)
        if pop != 0:
            pop = (vari(z[:4, :4]) + vari(z[:4, 4:]) + vari(z[4:, 4:]) + vari(z[4:, :4])) / pop
        return np.sqrt(m * pop) / 32

    out = np.zeros((nfactors, nby, nbx))
    for by in range(nby):
        for bx in range(nbx):
            za = A[4 * by:4 * by + 8, 4 * bx:4 * bx + 8]
            zb = B[4 * by:4 * by + 8, 4 * bx:4 * bx + 8]
            da, db = F.dctn(za, norm="ortho"), F.dctn(zb, norm="ortho")
            for f, (csf, mc, inv) in enumerate(tables):
                m = max(mask_of(za, da, mc), mask_of(zb, db, mc))
                e = np.maximum(np.abs(da - db) - m * inv, 0.0)
                out[f, by, bx] = ((e * csf) ** 2).sum()
    return out


def test_totals_match_float64_algorithm():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 256, (29, 37), dtype=np.uint8)
    b = np.clip(a.astype(int) + rng.integers(-20, 21, a.shape), 0, 255).astype(np.uint8)
    totals, tp = O.psnr_hvs_m_totals(a, b, 16)
    ref = _reference_f64(a, b, 16)
    nbx, nby = (37 - 1) // 4 + 1, (29 - 1) // 4 + 1
    assert tp == nbx * nby * 64
    exp = ref.sum(axis=(1, 2)) * 2.0 ** 24
    assert np.allclose(totals.astype(np.float64), exp, rtol=2e-4), (totals, exp)


def test_totals_properties():
    rng = np.random.default_rng(4)
    a = rng.integers(30, 220, (48, 64), dtype=np.uint8)
    t0, _ = O.psnr_hvs_m_totals(a, a)
    assert not t0.any()
    prev = None
    for amp in (2, 6, 18):
        b = np.clip(a.astype(int) + rng.integers(-amp, amp + 1, a.shape), 0, 255).astype(np.uint8)
        t, tp = O.psnr_hvs_m_totals(a, b)
        tr, _ = O.psnr_hvs_m_totals(b, a)
        assert np.array_equal(t, tr)
        if prev is not None:
            assert (t > prev).all()
        prev = t
        db = [O.psnr_hvs_m_db(x, tp) for x in t]
        assert all(np.isfinite(db))
    # dB of a known total: 10 log10(pixels * peak / error).
    assert abs(O.psnr_hvs_m_db(1 << 24, 1000, True) - 30.0) < 1e-12
    assert abs(O.psnr_hvs_m_db(1 << 24, 1000, False) - 10 * np.log10(1000 * (223.0 / 255) ** 2)) < 1e-6
