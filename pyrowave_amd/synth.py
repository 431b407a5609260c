"""Seeded synthetic YCbCr frames for tests and benchmarks.

Luma: a sum of 2-D linear gradients, plus band-limited Gaussian noise
(white noise blurred with an 8-pixel Gaussian, renormalised to sigma 12), plus
40 axis-aligned rectangles with uniform 0..255 fill and hard edges.  Chroma:
smooth gradients plus sigma-4 band-limited noise at chroma resolution.  All
planes are clamped to [0, 255] and rounded to uint8.  Optional 1-pixel
checker / text-like patches cover a given fraction of the luma area.

Sequences pan the scene by `pan` pixels per frame and redraw a fraction of the
rectangles each frame.
"""
import numpy as np

try:
    from scipy.ndimage import gaussian_filter as _gauss
except Exception:  # pragma: no cover - scipy is part of the image
    _gauss = None


def _blur(a, sigma):
    if _gauss is not None:
        return _gauss(a, sigma, mode="wrap")
    # FFT fallback with a Gaussian transfer function.
    h, w = a.shape
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    g = np.exp(-2.0 * (np.pi * sigma) ** 2 * (fx ** 2 + fy ** 2))
    return np.fft.irfft2(np.fft.rfft2(a) * g, s=a.shape)


def _noise(rng, h, w, sigma, corr):
    n = _blur(rng.standard_normal((h, w)).astype(np.float64), corr)
    s = n.std()
    return n * (sigma / s if s > 0 else 0.0)


def _gradients(rng, h, w, count, lo, hi):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    acc = np.zeros((h, w))
    for _ in range(count):
        ang = rng.uniform(0, 2 * np.pi)
        amp = rng.uniform(0.3, 1.0)
        d = (np.cos(ang) * xx / max(w, 1) + np.sin(ang) * yy / max(h, 1))
        acc += amp * d
    acc -= acc.min()
    if acc.max() > 0:
        acc /= acc.max()
    return lo + (hi - lo) * acc


def _rect(rng, h, w):
    rw = int(rng.integers(max(2, w // 50), max(3, w // 4)))
    rh = int(rng.integers(max(2, h // 50), max(3, h // 4)))
    x0 = int(rng.integers(0, max(1, w - rw)))
    y0 = int(rng.integers(0, max(1, h - rh)))
    return (y0, x0, rh, rw, float(rng.integers(0, 256)))


def _checker_patches(rng, h, w, fraction):
    """Mask and values of 1-px checker / text-like patches covering `fraction` of the area."""
    mask = np.zeros((h, w), bool)
    vals = np.zeros((h, w))
    target = fraction * h * w
    yy, xx = np.mgrid[0:h, 0:w]
    checker = ((xx + yy) & 1) * 255.0
    stripes = (xx & 1) * 255.0
    guard = 0
    while mask.sum() < target and guard < 10000:
        guard += 1
        ph = int(rng.integers(8, max(9, h // 8)))
        pw = int(rng.integers(8, max(9, w // 8)))
        y0 = int(rng.integers(0, max(1, h - ph)))
        x0 = int(rng.integers(0, max(1, w - pw)))
        kind = int(rng.integers(0, 3))
        sl = (slice(y0, y0 + ph), slice(x0, x0 + pw))
        if kind == 0:
            vals[sl] = checker[sl]
        elif kind == 1:
            vals[sl] = stripes[sl]
        else:  # glyph-like: random sparse 1-px strokes on a flat background
            v = rng.random((ph, pw)) < 0.3
            vals[sl] = np.where(v, 20.0, 235.0)
        mask[sl] = True
    return mask, vals


class SceneGenerator:
    """Generates frames of one seeded scene; frame t is panned by t * pan pixels."""

    def __init__(self, width, height, chroma420=True, seed=0, frames=1, pan=0, redraw=0.0, checker=0.0,
                 rects=40):
        self.width, self.height, self.chroma420 = width, height, chroma420
        self.rng = np.random.default_rng(seed)
        rng = self.rng
        self.pan = pan
        self.cw = width + pan * max(frames - 1, 0)
        self.redraw = redraw
        cw, h = self.cw, height
        base = _gradients(rng, h, cw, 3, 40.0, 215.0) + _noise(rng, h, cw, 12.0, 8.0)
        self.luma_bg = base
        self.rects = [_rect(rng, h, cw) for _ in range(rects)]
        if checker > 0:
            self.cmask, self.cvals = _checker_patches(rng, h, cw, checker)
        else:
            self.cmask = None
        ch_h, ch_w = (h // 2, cw // 2) if chroma420 else (h, cw)
        self.chroma_bg = [_gradients(rng, ch_h, ch_w, 2, 60.0, 196.0) + _noise(rng, ch_h, ch_w, 4.0, 8.0)
                          for _ in range(2)]

    def frame(self, t=0):
        """Returns [y, cb, cr] uint8 planes of frame t."""
        if t > 0 and self.redraw > 0:
            k = max(1, int(round(self.redraw * len(self.rects))))
            for _ in range(k):
                i = int(self.rng.integers(0, len(self.rects)))
                self.rects[i] = _rect(self.rng, self.height, self.cw)
        off = t * self.pan
        h, w = self.height, self.width
        y = self.luma_bg[:, off:off + w].copy()
        for (y0, x0, rh, rw, v) in self.rects:
            xs, xe = max(x0 - off, 0), min(x0 + rw - off, w)
            if xe > xs:
                y[y0:y0 + rh, xs:xe] = v
        if self.cmask is not None:
            m = self.cmask[:, off:off + w]
            y = np.where(m, self.cvals[:, off:off + w], y)
        planes = [y]
        if self.chroma420:
            co, cwd = off // 2, w // 2
        else:
            co, cwd = off, w
        for c in self.chroma_bg:
            planes.append(c[:, co:co + cwd])
        return [np.clip(np.rint(p), 0, 255).astype(np.uint8) for p in planes]


def frame(width, height, chroma420=True, seed=0, checker=0.0):
    """One seeded frame (config-0 style generator)."""
    return SceneGenerator(width, height, chroma420, seed=seed, checker=checker).frame(0)


def mirror_pattern(width, height, chroma420=True):
    """The reference's stability-test content (pyrowave_c_test.cpp:440-456)."""
    def mirror(v):
        v = v & 511
        return np.where(v > 255, 511 - v, v).astype(np.uint8)
    yy, xx = np.mgrid[0:height, 0:width]
    y = mirror(3 * xx + 5 * yy)
    ch, cw = (height // 2, width // 2) if chroma420 else (height, width)
    yy, xx = np.mgrid[0:ch, 0:cw]
    return [y, mirror(5 * yy + 7 * xx), mirror(7 * xx + 3 * yy)]


def psnr(a, b):
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 * 255.0 / mse)
