"""The two FP64 exponential forms of the Toda kernels
(nlchains_amd/csrc/nlc_exp.hpp: table-driven 2^(x/128) and table-free 2^y),
compiled on the host from the same header, against 2^(x/32) computed to 40
significant digits with Python's decimal module.  The tests state every
argument as x (1/32 octaves); the table form's wrapper passes 4x (1/128
octaves) and the table-free wrapper y = x/32, both exact, as the kernels'
state scaling does."""
import ctypes
import math
import re
import subprocess
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

HEADER = Path(__file__).resolve().parents[1] / "nlchains_amd" / "csrc" / "nlc_exp.hpp"

WRAPPER = r"""
#define NLC_EXP_QUAL inline
#define NLC_EXP_HOST 1
#include "nlc_exp.hpp"
static const double tab[nlc::kExpTabSize] = NLC_EXP2_TABLE;
extern "C" void exp_many(const double *y, double *out, long n) {
	for (long i = 0; i < n; ++i) out[i] = nlc::exp2_128_tab(y[i] * 4.0, tab);
}
extern "C" void exp_poly_many(const double *y, double *out, long n) {
	for (long i = 0; i < n; ++i) out[i] = nlc::exp2_poly(y[i] * 0x1p-5);
}
"""


@pytest.fixture(scope="module")
def exp_tab(tmp_path_factory):
    d = tmp_path_factory.mktemp("exp")
    src = d / "w.cpp"
    src.write_text(WRAPPER)
    so = d / "libw.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", f"-I{HEADER.parent}",
                    str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    funcs = {}
    for name in ("exp_many", "exp_poly_many"):
        fn = getattr(lib, name)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]

        def f(y, fn=fn):
            y = np.ascontiguousarray(y, dtype=np.float64)
            out = np.empty_like(y)
            fn(y.ctypes.data, out.ctypes.data, y.size)
            return out
        funcs[name] = f
    return funcs


@pytest.fixture(params=["exp_many", "exp_poly_many"])
def exp_fn(request, exp_tab):
    return exp_tab[request.param]


# Bound on the error in ulps: the rounded table entry (at most 1/2 ulp), the
# last FMA's rounding (1/2 ulp) and the polynomial (under 1/50 ulp) add up to
# just over 1 ulp where the two roundings meet; 60 000 random arguments per
# range reach 1.016 for the 256-entry table and 1.008 for the former 64-entry
# one.
ULP_BOUND = 1.02


def ulp_errors(x, got):
    """Error of got = 2^(x/32) in units of the last place of the exact value."""
    getcontext().prec = 40
    errs = []
    for xi, gi in zip(x, got):
        ref = Decimal(2) ** (Decimal(float(xi)) / 32)
        spacing = np.spacing(float(ref))
        errs.append(float(abs(Decimal(float(gi)) - ref) / Decimal(spacing)))
    return np.array(errs)


def test_table_entries_are_correctly_rounded():
    getcontext().prec = 60
    text = HEADER.read_text()
    block = text[text.index("#define NLC_EXP2_TABLE"):]
    block = block[:block.index("}") + 1]
    vals = [float.fromhex(h) for h in re.findall(r"0x[0-9a-f.]+p[+-]\d+", block)]
    assert len(vals) == 256
    for j, v in enumerate(vals):
        assert v == float(Decimal(2) ** (Decimal(j) / 256)), j


def annotated_literals(text, tag):
    """{k: literal} for the literals annotated `tag`k in comments of the header."""
    out = {}
    for line in text.splitlines():
        m = re.search(r"//\s*(%s\d+(?:,\s*%s\d+)*)\s*$" % (tag, tag), line)
        if not m:
            continue
        ks = [int(x) for x in re.findall(r"%s(\d+)" % tag, m.group(1))]
        lits = [float.fromhex(h) for h in re.findall(r"0x[0-9a-f.]+p[+-]\d+", line)]
        assert len(lits) == len(ks), line
        out.update(zip(ks, lits))
    return out


@pytest.mark.parametrize("tag,half,unit,ncoef,bound,taylor_miss", [
    ("p", 0.25, 1 / 128, 4, 5e-18, 3e-17),  # table form: 2^(r/128), |r| <= 1/4
    ("q", 0.5, 1, 11, 2.1e-17, 1e-15),       # table-free form: 2^r, |r| <= 1/2
])
def test_interpolants(tag, half, unit, ncoef, bound, taylor_miss):
    """r (c0 + ... + c_m r^m), with the literals annotated `tag`k in the header,
    is 2^(unit r) - 1 within `bound` on |r| <= half (60-digit check on a grid
    with the interval ends); the Taylor polynomial of the same degree misses by
    more than `taylor_miss`; and each coefficient, moved by about 2x the bound
    over the size of its term at r = half, fails."""
    getcontext().prec = 60
    coef = annotated_literals(HEADER.read_text(), tag)
    assert sorted(coef) == list(range(ncoef))
    a = Decimal(2).ln() * Decimal(unit)
    h = Decimal(half)

    def worst(c):
        w = Decimal(0)
        for i in range(-400, 401):
            r = h * i / 400
            acc = Decimal(0)
            for k in sorted(c, reverse=True):
                acc = acc * r + Decimal(c[k])
            w = max(w, abs(r * acc - ((r * a).exp() - 1)))
        return w

    assert worst(coef) < Decimal(bound)
    taylor = {k: float(a ** (k + 1) / math.factorial(k + 1)) for k in range(ncoef)}
    assert worst(taylor) > Decimal(taylor_miss)
    for k in range(ncoef):
        term = (half * math.log(2) * unit) ** (k + 1) / math.factorial(k + 1)
        bad = dict(coef)
        bad[k] = coef[k] * (1 + 10 ** math.ceil(math.log10(2 * bound / term)))
        assert worst(bad) > Decimal(bound), k


def test_shifter():
    """The rounding constant of the table form is 1.5 * 2^52."""
    text = HEADER.read_text()
    assert float.fromhex(re.search(r"kExpShifter = (0x[0-9a-f.]+p[+-]?\d+)", text).group(1)) == 1.5 * 2.0 ** 52


# x = 32 y / ln2 for y in (-1, 1), (-30, 30), (-700, 700), (-1e-6, 1e-6)
@pytest.mark.parametrize("lo,hi,count", [(-46.0, 46.0, 4000), (-1385.0, 1385.0, 4000), (-32300.0, 32300.0, 2000),
                                         (-4.6e-5, 4.6e-5, 500)])
def test_error_within_ulp_bound(exp_fn, lo, hi, count):
    x = np.random.default_rng(int(abs(hi) * 10)).uniform(lo, hi, count)
    err = ulp_errors(x, exp_fn(x))
    assert err.max() < ULP_BOUND, (err.max(), x[err.argmax()])


def test_reduction_boundaries_and_special_values(exp_fn):
    k = np.arange(-4000, 4000, dtype=np.float64)
    # interval ends of both reductions (N = nearest integer to 8x, the table
    # form's 2x' on x' = 4x in 1/128 octaves, resp. to x/32)
    h = k / 8
    x = np.concatenate([h, h + 0.0625, h - 0.0625, np.nextafter(h + 0.0625, 0), np.nextafter(h - 0.0625, 0),
                        k / 2 + 0.25, k / 2 - 0.25,
                        32 * k[::8] + 15.9999999, 32 * k[::8] - 15.9999999, 32 * k[::8] + 16.0])
    x = x[np.abs(x) < 32000]
    err = ulp_errors(x, exp_fn(x))
    assert err.max() < ULP_BOUND
    out = exp_fn(np.array([0.0, -0.0, 32767.0, 32768.0, 50000.0, -34500.0, -50000.0]))
    assert out[0] == 1.0 and out[1] == 1.0
    assert np.isfinite(out[2]) and out[3] == np.inf and out[4] == np.inf
    assert out[5] == 0.0 and out[6] == 0.0


def test_saturation_far_out_of_range(exp_fn):
    """Arguments far beyond the double range saturate: +inf above, 0 below, for
    every |x| < 2^50 (taking the integer part from the low bits of the shifted
    argument wrapped it from |x| = 2^30 on: 1.2e9 and 1e11 gave 0, some large
    negative x gave +inf).  The thresholds: 2^(x/32) is the largest finite
    double's neighbourhood at x = 32767.99 and under half the smallest
    subnormal below x = -34400.  The integer conversion of the host build
    saturates like the device's conversion instruction."""
    big = np.array([40000.0, 2.0 ** 30, 1.2e9, 1e9, 2.0 ** 31, 1e11, 2.0 ** 36, 1e15, 2.0 ** 49,
                    2.0 ** 50 - 1])
    assert np.all(exp_fn(np.concatenate([[32768.0, 33000.0], big])) == np.inf)
    assert np.all(exp_fn(np.concatenate([[-34401.0, -34500.0], -big])) == 0.0)
    near = np.array([32767.9, -34399.0, -34300.0])
    out = exp_fn(near)
    assert np.isfinite(out[0]) and out[0] > 1.7e308
    assert 0.0 < out[1] < 1e-323 and out[2] > 0.0
    x = 10.0 ** np.random.default_rng(7).uniform(5, np.log10(2.0 ** 50), 2000)
    assert np.all(exp_fn(x) == np.inf) and np.all(exp_fn(-x) == 0.0)


def test_beyond_exact_reduction_range(exp_tab):
    """From |4x| = 2^50 on, the table form's remainder 4x - N/2 is no longer
    exact: a positive argument still gives a non-finite value (the bond force
    of a blown-up chain is never finite) and a negative one 0 or a non-finite
    value, never a finite nonzero number.  The table-free form's remainder is
    exact for every finite x, so it saturates exactly."""
    x = 10.0 ** np.random.default_rng(8).uniform(np.log10(2.0 ** 50), 300, 4000)
    x = np.concatenate([x, [2.0 ** 52, 2.0 ** 60, 1e100, 1e300, np.finfo(np.float64).max]])
    exp_fn, exp_poly = exp_tab["exp_many"], exp_tab["exp_poly_many"]
    for out in (exp_fn(x), exp_poly(x)):
        assert not np.isfinite(out).any()
    neg = exp_fn(-x)
    assert np.all((neg == 0.0) | ~np.isfinite(neg))
    assert np.all(exp_poly(x) == np.inf) and np.all(exp_poly(-x) == 0.0)


def test_nonfinite_arguments_give_nan(exp_fn):
    """NaN and +-inf give NaN: such an argument only comes from a state that is
    already non-finite, and the run's non-finite-entropy check stops it."""
    assert np.isnan(exp_fn(np.array([np.nan, np.inf, -np.inf]))).all()
