"""The table-driven exp of the Toda kernels (nlchains_amd/csrc/nlc_exp.hpp),
compiled on the host from the same header, against e^y computed to 40
significant digits with Python's decimal module."""
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
    lib.exp_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]

    def f(y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(y)
        lib.exp_many(y.ctypes.data, out.ctypes.data, y.size)
        return out
    return f


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


@pytest.mark.parametrize("lo,hi,count", [(-1.0, 1.0, 4000), (-30.0, 30.0, 4000), (-700.0, 700.0, 2000),
                                         (-1e-6, 1e-6, 500)])
def test_error_within_one_ulp(exp_tab, lo, hi, count):
    y = np.random.default_rng(int(abs(hi) * 10)).uniform(lo, hi, count)
    err = ulp_errors(y, exp_tab(y))
    assert err.max() < 1.0, (err.max(), y[err.argmax()])


def test_table_boundaries_and_special_values(exp_tab):
    ln2 = np.log(2.0)
    k = np.arange(-2000, 2000)
    y = np.concatenate([k * ln2 / 32, k * ln2 / 32 + ln2 / 64, k * ln2 / 32 - ln2 / 64])
    err = ulp_errors(y, exp_tab(y))
    assert err.max() < 1.0
    out = exp_tab(np.array([0.0, -0.0, 709.7, 710.0, 1000.0, -746.0, -1000.0, np.nan]))
    assert out[0] == 1.0 and out[1] == 1.0
    assert np.isfinite(out[2]) and out[3] == np.inf and out[4] == np.inf
    assert out[5] == 0.0 and out[6] == 0.0 and np.isnan(out[7])
