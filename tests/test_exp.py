"""The two FP64 exp forms of the Toda kernels (nlchains_amd/csrc/nlc_exp.hpp:
table-driven and table-free), compiled on the host from the same header,
against e^y computed to 40 significant digits with Python's decimal module."""
import ctypes
import re
import subprocess
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

HEADER = Path(__file__).resolve().parents[1] / "nlchains_amd" / "csrc" / "nlc_exp.hpp"

WRAPPER = r"""
#define NLC_EXP_QUAL inline
#include "nlc_exp.hpp"
static const double tab[nlc::kExpTabSize] = NLC_EXP2_TABLE;
extern "C" void exp_many(const double *y, double *out, long n) {
	for (long i = 0; i < n; ++i) out[i] = nlc::exp_tab(y[i], tab);
}
extern "C" void exp_poly_many(const double *y, double *out, long n) {
	for (long i = 0; i < n; ++i) out[i] = nlc::exp_poly(y[i]);
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


def ulp_errors(y, got):
    getcontext().prec = 40
    errs = []
    for yi, gi in zip(y, got):
        ref = Decimal(float(yi)).exp()
        spacing = np.spacing(float(ref))
        errs.append(float(abs(Decimal(float(gi)) - ref) / Decimal(spacing)))
    return np.array(errs)


def test_table_entries_are_correctly_rounded():
    getcontext().prec = 60
    text = HEADER.read_text()
    block = text[text.index("#define NLC_EXP2_TABLE"):]
    block = block[:block.index("}") + 1]
    vals = [float.fromhex(h) for h in re.findall(r"0x[0-9a-f.]+p[+-]\d+", block)]
    assert len(vals) == 32
    for j, v in enumerate(vals):
        assert v == float(Decimal(2) ** (Decimal(j) / 32)), j


def test_reduction_constants():
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    text = HEADER.read_text()

    def const(name):
        return float.fromhex(re.search(name + r" = (0x[0-9a-f.]+p[+-]\d+)", text).group(1))
    hi, lo = const("kExpStepHi"), const("kExpStepLo")
    assert hi == float(ln2 / 32) and lo == float(ln2 / 32 - Decimal(hi))
    assert const("kExpInvStep") == float(32 / ln2)
    hi, lo = const("kExpLn2Hi"), const("kExpLn2Lo")
    assert hi == float(ln2) and lo == float(ln2 - Decimal(hi))
    assert const("kExpInvLn2") == float(1 / ln2)


def test_polynomial_coefficients_are_reciprocal_factorials():
    """Every literal annotated 1/k! in the header is 1/k! rounded to nearest."""
    getcontext().prec = 60
    fact = [Decimal(1)]
    for k in range(1, 20):
        fact.append(fact[-1] * k)
    checked = 0
    for line in HEADER.read_text().splitlines():
        m = re.search(r"//\s*((?:1/\d+!?,?\s*)+)$", line)
        if not m:
            continue
        ks = [int(x) for x in re.findall(r"1/(\d+)", m.group(1))]
        lits = [float.fromhex(h) for h in re.findall(r"0x[0-9a-f.]+p[+-]\d+", line)]
        # a "1/n" without "!" is the value itself (1/720 = 1/6!, 1/120 = 1/5!, ...)
        for k, lit in zip(ks, lits):
            want = 1 / fact[k] if f"1/{k}!" in m.group(1) else 1 / Decimal(k)
            assert lit == float(want), (line, k)
            checked += 1
    assert checked >= 15


@pytest.mark.parametrize("lo,hi,count", [(-1.0, 1.0, 4000), (-30.0, 30.0, 4000), (-700.0, 700.0, 2000),
                                         (-1e-6, 1e-6, 500)])
def test_error_within_one_ulp(exp_fn, lo, hi, count):
    y = np.random.default_rng(int(abs(hi) * 10)).uniform(lo, hi, count)
    err = ulp_errors(y, exp_fn(y))
    assert err.max() < 1.0, (err.max(), y[err.argmax()])


def test_reduction_boundaries_and_special_values(exp_fn):
    ln2 = np.log(2.0)
    k = np.arange(-2000, 2000)
    # interval ends of both reductions (steps ln2/32 and ln2)
    y = np.concatenate([k * ln2 / 32, k * ln2 / 32 + ln2 / 64, k * ln2 / 32 - ln2 / 64,
                        k[::4] * ln2 + 0.4999999 * ln2, k[::4] * ln2 - 0.4999999 * ln2])
    y = y[np.abs(y) < 700]
    err = ulp_errors(y, exp_fn(y))
    assert err.max() < 1.0
    out = exp_fn(np.array([0.0, -0.0, 709.7, 710.0, 1000.0, -746.0, -1000.0, np.nan]))
    assert out[0] == 1.0 and out[1] == 1.0
    assert np.isfinite(out[2]) and out[3] == np.inf and out[4] == np.inf
    assert out[5] == 0.0 and out[6] == 0.0 and np.isnan(out[7])
