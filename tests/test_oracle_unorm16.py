"""16-bit (R16_UNORM) input and output of the oracle (CPU).

The reference encodes 16-bit y4m input through R16_UNORM images
(encode.cpp:152, yuv4mpeg.cpp:221-233 rescales p10-p16 to the full 16-bit
range) and decodes into R16_UNORM targets (decode.cpp:215).  Pinned here:

- a 16-bit sample 257 * u is the rational u / 255, so planes of 257 * u8
  give the 8-bit transform and bitstream exactly, in every precision and in
  FP64 (closed form);
- the level-0 conversion is fl(fl(u / 65535) - 1/2) (FP32) and the fp16
  chain of precision 0, against exact rationals, for every 8-bit value and
  every seventh 16-bit value with the ends of the range;
- the 16-bit FP32 transform of full-range 16-bit content is within the FP32
  bound of the FP64 transform;
- 16-bit decode output is round(clamp(v + 0.5) * 65535) of the FP64 inverse
  within one unit, and its 8-bit reduction agrees with the 8-bit decode.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from pyrowave_amd import synth


def _planes16(w, h, chroma420, seed):
    rng = np.random.default_rng(seed)
    base = synth.frame(w, h, chroma420, seed=seed)
    return [(p.astype(np.uint32) * 257 + rng.integers(-200, 201, p.shape)).clip(0, 65535).astype(np.uint16)
            for p in base]


@pytest.mark.parametrize("w,h,chroma,precision", [(64, 48, 0, 2), (70, 38, 1, 1), (96, 64, 0, 3), (66, 34, 0, 0)])
def test_257_times_u8_equals_u8(w, h, chroma, precision):
    p8 = synth.frame(w, h, chroma == 0, seed=w + h)
    p16 = [p.astype(np.uint16) * 257 for p in p8]
    L = O.layout(w, h, chroma, precision=0 if precision == 0 else None)
    a = O.dwt_frame(L, p8, precision=precision)
    b = O.dwt_frame(L, p16, precision=precision)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(O.dwt_frame(L, p8, f64=True), O.dwt_frame(L, p16, f64=True))
    budget = w * h // 3
    bs8, meta8, _ = O.encode(L, p8, budget, precision=precision)
    bs16, meta16, _ = O.encode(L, p16, budget, precision=precision)
    assert np.array_equal(bs8, bs16) and np.array_equal(meta8, meta16)


def _round_f32(x):
    """Exact round-to-nearest-even of a Fraction to float32 (normal range)."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    scale = Fraction(2) ** (23 - e)
    n = a * scale
    fl = n.numerator // n.denominator
    rem = n - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2):
        fl += 1
    return sign * Fraction(fl) / scale


def _round_f16(x):
    """Exact round-to-nearest-even of a Fraction to fp16 (normal range,
    |x| >= 2^-14, or subnormal spacing 2^-24 below it)."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -14)
    scale = Fraction(2) ** (10 - e)
    n = a * scale
    fl = n.numerator // n.denominator
    rem = n - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2):
        fl += 1
    return sign * Fraction(fl) / scale


@pytest.mark.parametrize("bits", [8, 16])
def test_level0_conversion_exact(bits):
    """The oracle's level-0 sample is fl(fl(u / (2^bits - 1)) - 1/2) for every
    sample value (FP32), and fp16(fp16(fl(u / (2^bits - 1))) - 1/2) in
    precision 0; computed here with exact rationals."""
    n = (1 << bits) - 1
    step = 1 if bits == 8 else 7
    values = sorted(set(list(range(0, n + 1, step)) + [1, 2, n // 2, n // 2 + 1, n - 1, n]))
    for u in values:
        t = _round_f32(Fraction(u, n))
        assert Fraction(O.unorm_dc(u, bits, 2)) == _round_f32(t - Fraction(1, 2)), (bits, u)
        assert Fraction(O.unorm_dc(u, bits, 0)) == _round_f16(_round_f16(t) - Fraction(1, 2)), (bits, u)


def test_fp32_transform_of_16bit_content_vs_fp64():
    w, h = 130, 66
    p16 = _planes16(w, h, True, seed=3)
    L = O.layout(w, h, 0)
    c32 = O.dwt_frame(L, p16, precision=2).astype(np.float64)
    c64 = O.dwt_frame(L, p16, f64=True)
    assert np.max(np.abs(c32 - c64)) <= 2e-5
    # The 16-bit detail is visible: the transform differs from the 8-bit one.
    p8 = [(p.astype(np.uint32) + 128) // 257 for p in p16]
    c8 = O.dwt_frame(L, [p.astype(np.uint8) for p in p8], f64=True)
    assert np.max(np.abs(c8 - c64)) > 1e-4


def test_16bit_decode_output():
    w, h = 96, 64
    p16 = _planes16(w, h, True, seed=9)
    L = O.layout(w, h, 0)
    coeffs = O.dwt_frame(L, p16, precision=2)
    out16 = O.idwt_frame(L, coeffs, precision=2, bits=16)
    out8 = O.idwt_frame(L, coeffs, precision=2)
    ref64 = O.idwt_frame_f64(L, coeffs.astype(np.float64))
    for c in range(3):
        exp = np.floor(np.clip(ref64[c], 0.0, 1.0) * 65535.0 + 0.5)  # ref64 carries the +0.5 shift
        assert out16[c].dtype == np.uint16
        assert np.max(np.abs(out16[c].astype(np.float64) - exp)) <= 1
        # The unquantized round trip returns the 16-bit input within a unit.
        assert np.max(np.abs(out16[c].astype(np.int64) - p16[c].astype(np.int64))) <= 1
        # Reduced to 8 bits it agrees with the 8-bit store within one unit.
        assert np.max(np.abs(np.round(out16[c] / 257.0) - out8[c].astype(np.float64))) <= 1
    # Clamp and rounding at the ends of the range.
    zero = np.zeros(L.coeff_count, np.float32)
    mid = O.idwt_frame(L, zero, precision=2, bits=16)
    assert all(np.all(p == 32768) for p in mid)


def test_16bit_decode_of_encoded_stream():
    """Encode 16-bit content, decode into 16-bit planes: PSNR at 16 bits grows
    with the budget."""
    w, h = 128, 96
    p16 = _planes16(w, h, True, seed=21)
    L = O.layout(w, h, 0)
    vals = []
    for budget in (2000, 6000, 20000):
        bs, meta, _ = O.encode(L, p16, budget, precision=2)
        out, pk = O.packetize(L, meta, bs, 1, 1 << 30)
        dec = O.decode(L, [out[o:o + s].tobytes() for o, s in pk], precision=2, bits=16)
        err = dec[0].astype(np.float64) - p16[0]
        vals.append(10 * np.log10(65535.0 ** 2 / max(np.mean(err * err), 1e-12)))
    assert vals[0] < vals[1] < vals[2], vals
