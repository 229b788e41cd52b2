"""ctypes front-end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
DNKG, FPUT, TODA, DDNKG, DNLS = range(5)

_dp = C.POINTER(C.c_double)


def _load(name: str) -> C.CDLL:
    path = _HERE / name
    if not path.exists():
        raise ImportError(f"{path} is not built: run oracle/build.sh")
    lib = C.CDLL(str(path))
    sig = {
        "orc_yoshida": (None, [_dp, _dp]),
        "orc_coeffs": (None, [C.c_double, _dp, _dp]),
        "orc_move_real": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, _dp, C.c_uint32, C.c_double, C.c_double,
                                    C.c_double, C.c_double, _dp, C.c_int]),
        "orc_forces": (C.c_int, [C.c_int, C.c_uint32, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
        "orc_fft": (None, [C.c_uint32, _dp, _dp, C.c_int]),
        "orc_dispersion_real": (None, [C.c_uint32, C.c_double, _dp]),
        "orc_dispersion_dnls": (None, [C.c_uint32, _dp]),
        "orc_linenergies_real": (None, [C.c_uint32, C.c_uint32, _dp, _dp, _dp]),
        "orc_calc_linenergies": (None, [C.c_uint32, _dp, C.c_double, C.c_int]),
        "orc_entropy_modes": (C.c_int, [C.c_uint32, C.c_void_p, C.c_int, C.POINTER(C.c_uint32)]),
        "orc_entropies": (None, [C.c_uint32, _dp, C.c_int, C.POINTER(C.c_uint32), C.c_uint32, _dp, _dp]),
        "orc_eigensystem": (C.c_int, [C.c_uint32, _dp, _dp, _dp]),
        "orc_linenergies_ddnkg": (None, [C.c_uint32, C.c_uint32, _dp, _dp, _dp, _dp]),
        "orc_dnls_tables": (None, [C.c_uint32, C.c_double, _dp]),
        "orc_move_dnls": (C.c_int, [C.c_uint32, C.c_uint32, _dp, C.c_uint32, C.c_double, C.c_double, C.c_int]),
        "orc_linenergies_dnls": (None, [C.c_uint32, C.c_uint32, _dp, _dp]),
        "orc_chain_energy": (None, [C.c_int, C.c_uint32, C.c_uint32, _dp, C.c_double, C.c_double, C.c_double,
                                    _dp, _dp]),
        "orc_dnls_norm": (None, [C.c_uint32, C.c_uint32, _dp, _dp]),
        "orc_has_openmp": (C.c_int, []),
    }
    for name_, (res, args) in sig.items():
        f = getattr(lib, name_)
        f.restype, f.argtypes = res, args
    return lib


_strict = None
_fast = None


def strict() -> C.CDLL:
    global _strict
    if _strict is None:
        _strict = _load("liboracle.so")
    return _strict


def fast() -> C.CDLL:
    global _fast
    if _fast is None:
        _fast = _load("liboracle_fast.so")
    return _fast


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def yoshida():
    c, d = np.zeros(8), np.zeros(8)
    strict().orc_yoshida(_p(c), _p(d))
    return c, d


def move_real(model: int, state: np.ndarray, steps: int, dt: float, m=0.0, alpha=0.0, beta=0.0, mp2=None,
              threads: int = 1, lib: C.CDLL | None = None) -> np.ndarray:
    """Returns a new array: state advanced by `steps` Yoshida-6 steps (planar layout)."""
    lib = lib or strict()
    s = np.array(state, dtype=np.float64, order="C", copy=True)
    copies, n, _ = s.shape
    mp = None if mp2 is None else np.ascontiguousarray(mp2, dtype=np.float64)
    rc = lib.orc_move_real(model, n, copies, _p(s), steps, dt, m, alpha, beta, None if mp is None else _p(mp),
                           threads)
    if rc:
        raise ValueError("orc_move_real rejected its arguments")
    return s


def forces(model: int, q: np.ndarray, m=0.0, alpha=0.0, beta=0.0, mp2=None) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.zeros_like(q)
    mp = None if mp2 is None else np.ascontiguousarray(mp2, dtype=np.float64)
    strict().orc_forces(model, q.size, _p(q), m, alpha, beta, None if mp is None else _p(mp), _p(out))
    return out


def fft(z: np.ndarray, sign: int = -1) -> np.ndarray:
    re = np.ascontiguousarray(z.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(z.imag, dtype=np.float64).copy()
    strict().orc_fft(re.size, _p(re), _p(im), sign)
    return re + 1j * im


def dispersion_real(n: int, m: float) -> np.ndarray:
    w = np.zeros(n)
    strict().orc_dispersion_real(n, m, _p(w))
    return w


def dispersion_dnls(n: int) -> np.ndarray:
    w = np.zeros(n)
    strict().orc_dispersion_dnls(n, _p(w))
    return w


def linenergies_real(state: np.ndarray, omega: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    copies, n, _ = s.shape
    e = np.zeros(n)
    strict().orc_linenergies_real(n, copies, _p(s), _p(np.ascontiguousarray(omega, dtype=np.float64)), _p(e))
    return e


def calc_linenergies(e_raw: np.ndarray, norm: float, a0is0: bool) -> np.ndarray:
    e = np.array(e_raw, dtype=np.float64, copy=True)
    strict().orc_calc_linenergies(e.size, _p(e), norm, int(a0is0))
    return e


def entropy_modes(n: int, mask, a0is0: bool):
    """Mode index list per the --entropymask rules, [] for no mask, None if too few modes."""
    if mask is None:
        return []
    mk = np.ascontiguousarray(mask, dtype=np.uint8)
    idx = (C.c_uint32 * n)()
    cnt = strict().orc_entropy_modes(n, mk.ctypes.data_as(C.c_void_p), int(a0is0), idx)
    return None if cnt < 0 else list(idx[:cnt])


def entropies(e: np.ndarray, a0is0: bool, idx=()) -> tuple[float, float]:
    e = np.ascontiguousarray(e, dtype=np.float64)
    wt, inf = C.c_double(), C.c_double()
    arr = (C.c_uint32 * max(1, len(idx)))(*idx)
    strict().orc_entropies(e.size, _p(e), int(a0is0), arr, len(idx), C.byref(wt), C.byref(inf))
    return wt.value, inf.value


def eigensystem(mp2: np.ndarray):
    """(omegas ascending, V with V[:, j] = mode j, sweeps)."""
    mp2 = np.ascontiguousarray(mp2, dtype=np.float64)
    n = mp2.size
    om, vt = np.zeros(n), np.zeros((n, n))
    sweeps = strict().orc_eigensystem(n, _p(mp2), _p(om), _p(vt))
    return om, vt.T.copy(), sweeps


def linenergies_ddnkg(state: np.ndarray, V: np.ndarray, omega: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    copies, n, _ = s.shape
    vt = np.ascontiguousarray(V.T)
    e = np.zeros(n)
    strict().orc_linenergies_ddnkg(n, copies, _p(s), _p(vt), _p(np.ascontiguousarray(omega, dtype=np.float64)),
                                   _p(e))
    return e


def dnls_tables(n: int, dt: float) -> np.ndarray:
    t = np.zeros((7, n, 2))
    strict().orc_dnls_tables(n, dt, _p(t))
    return t[..., 0] + 1j * t[..., 1]


def move_dnls(state: np.ndarray, steps: int, dt: float, beta: float, threads: int = 1,
              lib: C.CDLL | None = None) -> np.ndarray:
    lib = lib or strict()
    s = np.array(state, dtype=np.float64, order="C", copy=True)
    copies, n, _ = s.shape
    if lib.orc_move_dnls(n, copies, _p(s), steps, dt, beta, threads):
        raise ValueError("orc_move_dnls rejected its arguments")
    return s


def linenergies_dnls(state: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    copies, n, _ = s.shape
    e = np.zeros(n)
    strict().orc_linenergies_dnls(n, copies, _p(s), _p(e))
    return e


def chain_energy(model: int, state: np.ndarray, m=0.0, alpha=0.0, beta=0.0, masses=None) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    copies, n, _ = s.shape
    E = np.zeros(copies)
    ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    strict().orc_chain_energy(model, n, copies, _p(s), m, alpha, beta, None if ms is None else _p(ms), _p(E))
    return E


def dnls_norm(state: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    copies, n, _ = s.shape
    out = np.zeros(copies)
    strict().orc_dnls_norm(n, copies, _p(s), _p(out))
    return out
