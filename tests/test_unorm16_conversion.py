"""The 16-bit level-0 conversion of the HIP kernels (CPU).

The transform kernels turn a 16-bit sample u into float(u / 65535) without a
division: with the float pair hi = 0x37800080, lo = 0x27800080 (hi + lo
approximates 1/65535), fmaf(u, hi, u * lo).  Checked here exhaustively, with
exact rationals, to round to the correctly rounded quotient for every u in
0..65535, as the 8-bit conversion is for 0..255 with its own pair.  The GPU
tests then compare the kernels' coefficients with the oracle on a plane that
holds every 16-bit value.
"""
import struct
from fractions import Fraction

import pytest


def _f32(x):
    """Exact round-to-nearest-even of a Fraction to float32 (normal range)."""
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    sc = Fraction(2) ** (23 - e)
    n = a * sc
    fl = n.numerator // n.denominator
    r = n - fl
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2):
        fl += 1
    return s * Fraction(fl) / sc


def _bits(u32):
    return Fraction(struct.unpack(">f", struct.pack(">I", u32))[0])


@pytest.mark.parametrize("n,hi,lo,step", [(65535, 0x37800080, 0x27800080, 1), (255, 0x3B808081, 0xAF7EFEFF, 1)])
def test_split_constant_quotient_is_correctly_rounded(n, hi, lo, step):
    H, Lo = _bits(hi), _bits(lo)
    assert abs(H + Lo - Fraction(1, n)) < Fraction(1, n) * Fraction(1, 2 ** 40)
    for u in range(0, n + 1, step):
        fu = Fraction(u)
        got = _f32(fu * H + _f32(fu * Lo))  # fmaf: one rounding of the exact sum
        assert got == _f32(Fraction(u, n)), u
