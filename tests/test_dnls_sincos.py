"""The constants of the DNLS nonlinear rotation e^{-i theta |psi|^2}
(nlc_dnls_common.inl: expi_small, kSixteenth, expi_reduced, kExpiSmall),
read from the kernel source and checked against sin and cos to 60 digits.

The reference takes the rotation from the library sincos (DNLS.cu:11-17);
the kernels evaluate it as Taylor polynomials of degree 9 (sin) and 10 (cos)
on |r| <= pi/32 after a three-part Cody & Waite reduction by pi/16.  What
fixes these constants is the mathematics: the polynomial, evaluated exactly,
must stay within 3e-18 (relative) of sin and cos on the whole interval, the
table entries must be the correctly rounded e^{i k pi/16}, the three parts of
pi/16 must sum to it, and the short form's bound must give k = 0 in the
reduced form (so that the two forms agree bit for bit below it).  A dropped
or mistyped coefficient, a wrong table entry or a wrong split fails here.
"""
from __future__ import annotations

import re
from decimal import Decimal, getcontext
from fractions import Fraction
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parents[1] / "nlchains_amd" / "csrc" / "nlc_dnls_common.inl"

getcontext().prec = 60
PI = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")
NUM = r"(-?0x[0-9a-fA-F.]+p-?\d+|-?\d+\.\d+(?:e-?\d+)?)"


def _num(s: str) -> float:
    return float.fromhex(s) if "0x" in s else float(s)


def _series(x: Decimal, first: Decimal, n0: int) -> Decimal:
    total, term, n = Decimal(0), first, n0
    while abs(term) > Decimal(10) ** -58:
        total += term
        term = -term * x * x / ((n + 1) * (n + 2))
        n += 2
    return total


def dsin(x: Decimal) -> Decimal:
    return _series(x, x, 1)


def dcos(x: Decimal) -> Decimal:
    return _series(x, Decimal(1), 0)


@pytest.fixture(scope="module")
def src() -> str:
    return SRC.read_text()


def _body(src: str, name: str) -> str:
    start = src.index(name)
    return src[start:src.index("\n}\n", start)]


@pytest.fixture(scope="module")
def poly(src):
    """Horner coefficients of expi_small, highest degree first."""
    body = _body(src, "double2 expi_small(double r)")
    s = [_num(m) for m in re.findall(r"double s = " + NUM, body) + re.findall(r"s = fma\(s, x, " + NUM, body)]
    c = [_num(m) for m in re.findall(r"double c = " + NUM, body) + re.findall(r"c = fma\(c, x, " + NUM, body)]
    # the returns: cos = fma(c, x, 1.0), sin = fma(r * x, s, r)
    assert "make_double2(fma(c, x, 1.0), fma(r * x, s, r))" in body
    return s, c


def _eval(s, c, r: float):
    """(cos, sin) of the kernel's polynomial, evaluated exactly."""
    x = Fraction(r) ** 2
    ps = Fraction(0)
    for k in s:
        ps = ps * x + Fraction(k)
    pc = Fraction(0)
    for k in c:
        pc = pc * x + Fraction(k)
    return pc * x + 1, Fraction(r) * x * ps + Fraction(r)


def test_polynomial_degrees(poly):
    s, c = poly
    assert len(s) == 4 and len(c) == 5  # sin to r^9, cos to r^10


def test_polynomial_within_3e18_of_sin_and_cos(poly):
    s, c = poly
    half = float(PI / 32) * (1 + 1e-12)  # reduced arguments reach pi/32 and a rounding beyond
    worst_s = worst_c = 0.0
    for i in range(-60, 61):
        r = half * i / 60
        if r == 0:
            continue
        pc, ps = _eval(s, c, r)
        rd = Decimal(r)
        es = abs(Decimal(ps.numerator) / Decimal(ps.denominator) - dsin(rd)) / abs(rd)
        ec = abs(Decimal(pc.numerator) / Decimal(pc.denominator) - dcos(rd))
        worst_s, worst_c = max(worst_s, float(es)), max(worst_c, float(ec))
    assert worst_s < 3e-18 and worst_c < 3e-18, (worst_s, worst_c)
    # one degree less would not do: the first dropped sin term alone is 1e-16
    assert float((PI / 32) ** 8 / 362880) > 1e-17


def test_table_is_correctly_rounded(src):
    body = src[src.index("kSixteenth[32] = {"):]
    body = body[:body.index("};")]
    pairs = re.findall(r"\{" + NUM + r", " + NUM + r"\}", body)
    assert len(pairs) == 32
    def rounded(v: Decimal) -> float:  # the exact zeros of the quarter turns come out near 1e-60
        return float(v) + 0.0 if abs(v) > Decimal(10) ** -50 else 0.0

    for k, (cs, sn) in enumerate(pairs):
        a = PI * k / 16
        assert _num(cs) == rounded(dcos(a)), k
        assert _num(sn) == rounded(dsin(a)), k


def test_reduction_constants(src):
    body = _body(src, "double2 expi_reduced(double a)")
    q = {n: _num(v) for n, v in re.findall(r"(kQ[123]) = " + NUM, body)}
    inv = _num(re.search(r"kSixteenOverPi = " + NUM, body).group(1))
    assert "kSixteenth[int(k) & 31]" in body
    total = Decimal(q["kQ1"]) + Decimal(q["kQ2"]) + Decimal(q["kQ3"])
    assert abs(total - PI / 16) < Decimal(10) ** -36
    # k * Q1 and k * Q2 exact for |k| < 2^20: at most 33 significant bits
    for n in ("kQ1", "kQ2"):
        m = Fraction(q[n])
        while m.denominator != 1:
            m *= 2
        assert m.numerator.bit_length() - (m.numerator & -m.numerator).bit_length() + 1 <= 33, n
    assert inv == float(16 / PI)
    small = _num(re.search(r"constexpr double kExpiSmall = " + NUM, src).group(1))
    assert small <= float(PI / 32)
    assert small * inv < 0.5  # every angle the short form takes has k = 0
