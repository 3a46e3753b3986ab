"""Exact posit32 (es=2) semantics from the definition, with Python rationals.

Independent of the oracle: a pattern is decoded by reading its bit string (sign,
regime run, two exponent bits, fraction), and a rational is encoded by writing the
exact infinite bit string and rounding it to 31 bits after the sign, to nearest with
ties to the even pattern, saturating to minpos / maxpos (a nonzero value never becomes
0 or NaR).  Used to pin the oracle's div, sqrt and double conversion everywhere, and
its add and mul wherever the result's regime is at most 28 bits long.
"""
from fractions import Fraction

NAR = 0x80000000
MASK = 0xFFFFFFFF


def decode(p):
    """Pattern -> Fraction (None for NaR)."""
    p &= MASK
    if p == 0:
        return Fraction(0)
    if p == NAR:
        return None
    neg = p >> 31
    if neg:
        p = (-p) & MASK
    bits = format(p, "032b")[1:]
    first = bits[0]
    run = len(bits) - len(bits.lstrip(first))
    k = run - 1 if first == "1" else -run
    rest = bits[run + 1:]
    e = int((rest[:2] + "00")[:2], 2)
    fb = rest[2:]
    f = Fraction(int(fb, 2), 1 << len(fb)) if fb else Fraction(0)
    v = (1 + f) * Fraction(2) ** (4 * k + e)
    return -v if neg else v


def regime_len(p):
    """Length of the regime run of a finite nonzero pattern."""
    p &= MASK
    if p >> 31:
        p = (-p) & MASK
    bits = format(p, "032b")[1:]
    return len(bits) - len(bits.lstrip(bits[0]))


def scale_of(x):
    """floor(log2(|x|)) for a nonzero Fraction."""
    x = abs(x)
    s = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** s > x:
        s -= 1
    return s


def encode(x):
    """Fraction -> pattern, RNE on the posit bit string, saturating."""
    if x is None:
        return NAR
    if x == 0:
        return 0
    neg = x < 0
    x = abs(x)
    s = scale_of(x)
    k, e = divmod(s, 4)
    if k >= 30:
        p = 0x7FFFFFFF
    elif k <= -31:
        p = 1
    else:
        r = -k if k < 0 else k + 1
        regime = ("1" * r + "0") if k >= 0 else ("0" * r + "1")
        head = regime + format(e, "02b")
        nf = max(0, 33 - len(head)) + 8
        m = x / Fraction(2) ** s - 1  # in [0, 1)
        scaled = m * (1 << nf)
        fint = scaled.numerator // scaled.denominator
        exact = scaled == fint
        s_bits = head + format(fint, "0%db" % nf)
        kept = int(s_bits[:31], 2)
        guard = s_bits[31] == "1"
        sticky = ("1" in s_bits[32:]) or not exact
        if guard and (sticky or (kept & 1)):
            kept += 1
        p = kept
    return ((-p) & MASK) if neg else p


def add(a, b):
    x, y = decode(a), decode(b)
    if x is None or y is None:
        return NAR
    return encode(x + y)


def mul(a, b):
    x, y = decode(a), decode(b)
    if x is None or y is None:
        return NAR
    return encode(x * y)


def div(a, b):
    x, y = decode(a), decode(b)
    if x is None or y is None or y == 0:
        return NAR
    return encode(x / y)


def sqrt(a):
    """Correctly rounded square root via integer square roots of the exact value."""
    x = decode(a)
    if x is None or x < 0:
        return NAR
    if x == 0:
        return 0
    # sqrt(x) lies in [2^h, 2^(h+1)) with h = floor(scale/2); take 80 bits beyond.
    h = scale_of(x) // 2
    q = 80
    # t = floor(sqrt(x) * 2^(q-h)) computed as isqrt(floor(x * 2^(2(q-h))))
    y = x * Fraction(2) ** (2 * (q - h))
    yi = y.numerator // y.denominator
    t = _isqrt(yi)
    exact = (t * t == y)
    val = Fraction(t, 1) / Fraction(2) ** (q - h)
    if exact:
        return encode(val)
    # nudge by a quarter ulp of the 80-bit window so the sticky bit is seen
    return encode(val + Fraction(1, 4) / Fraction(2) ** (q - h))


def _isqrt(n):
    if n < 2:
        return n
    x = 1 << ((n.bit_length() + 1) // 2)
    while True:
        y = (x + n // x) // 2
        if y >= x:
            return x
        x = y


def from_double(d):
    import math

    if d == 0.0:
        return 0
    if math.isnan(d) or math.isinf(d):
        return NAR
    return encode(Fraction(d))
