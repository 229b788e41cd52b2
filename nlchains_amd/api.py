"""ctypes binding of libnlchains_amd.so (include/nlc.h).

The hot path runs only on the GPU through the C ABI; there is no CPU fallback.
Importing this module fails loudly when the shared library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("NLC_LIBRARY", _PKG / "libnlchains_amd.so"))

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built: run `python -m nlchains_amd.build` (or __graft_entry__.build())")

lib = C.CDLL(str(LIB_PATH))

DNKG, FPUT, TODA, DDNKG, DNLS = range(5)
MODELS = {"DNKG": DNKG, "FPUT": FPUT, "Toda": TODA, "dDNKG": DDNKG, "DNLS": DNLS}
FP64, FP32C = 0, 1
KERNEL_AUTO, KERNEL_WAVE, KERNEL_SPLIT, KERNEL_TILED = 0, 1, 2, 3
PROJ_AUTO, PROJ_MFMA_F64, PROJ_VALU_F64, PROJ_SPLIT_BF16, PROJ_MFMA_F64_WIDE = 0, 1, 2, 3, 4

NLC_OK, NLC_EARG, NLC_EHIP, NLC_ENOMEM, NLC_EUNSUPPORTED, NLC_ESTATE = 0, -1, -2, -3, -4, -5


class Params(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("model", C.c_int32),
        ("precision", C.c_int32),
        ("kernel", C.c_int32),
        ("chain_length", C.c_uint32),
        ("copies", C.c_uint32),
        ("dt", C.c_double),
        ("m", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("masses", C.POINTER(C.c_double)),
        ("lanes_per_chain", C.c_uint32),
        ("sites_per_lane", C.c_uint32),
        ("projection", C.c_int32),
    ]


_dp = C.POINTER(C.c_double)
_ctx_p = C.c_void_p


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("nlc_abi_version", C.c_int)
_sig("nlc_strerror", C.c_char_p, C.c_int)
_sig("nlc_last_error", C.c_char_p, _ctx_p)
_sig("nlc_params_init", C.c_int, C.POINTER(Params))
_sig("nlc_create", C.c_int, C.POINTER(_ctx_p), C.POINTER(Params))
_sig("nlc_destroy", None, _ctx_p)
_sig("nlc_upload", C.c_int, _ctx_p, C.c_void_p)
_sig("nlc_download", C.c_int, _ctx_p, C.c_void_p)
_sig("nlc_device_state", C.c_int, _ctx_p, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t))
_sig("nlc_move", C.c_int, _ctx_p, C.c_uint32)
_sig("nlc_set_entropy_mask", C.c_int, _ctx_p, C.c_void_p, C.c_int)
_sig("nlc_linenergies", C.c_int, _ctx_p, C.c_double, _dp, _dp)
_sig("nlc_linenergies_raw", C.c_int, _ctx_p, _dp)
_sig("nlc_chain_energies", C.c_int, _ctx_p, _dp, _dp)
_sig("nlc_sync", C.c_int, _ctx_p)
_sig("nlc_stream", C.c_void_p, _ctx_p)
_sig("nlc_describe", C.c_int, _ctx_p, C.c_char_p, C.c_size_t)
_sig("nlc_yoshida", C.c_int, _dp, _dp)
_sig("nlc_dispersion", C.c_int, C.c_int, C.c_uint32, C.c_double, _dp)
_sig("nlc_eigensystem", C.c_int, C.c_uint32, _dp, _dp, _dp)
_sig("nlc_ic_plane_waves", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double,
     C.c_uint64, _dp)
_sig("nlc_ic_dnls", C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, _dp)
_sig("nlc_ic_eigenmodes", C.c_int, C.c_uint32, C.c_uint32, _dp, _dp, C.c_uint32, C.c_double, C.c_uint64, _dp)
_sig("nlc_ic_masses", C.c_int, C.c_uint32, C.c_double, C.c_double, C.c_uint64, _dp)
_u32p = C.POINTER(C.c_uint32)
_sig("nlc_create_ragged", C.c_int, C.POINTER(_ctx_p), C.POINTER(Params), _u32p, C.c_uint32)
_sig("nlc_ragged_classes", C.c_int, _ctx_p, _u32p, _u32p, _u32p, C.c_uint32)
_sig("nlc_create_mixed", C.c_int, C.POINTER(_ctx_p), C.POINTER(Params), C.c_uint32, _u32p, _u32p, C.c_uint32)
_sig("nlc_ragged_class_sets", C.c_int, _ctx_p, _u32p, C.c_uint32)
_sig("nlc_batch", C.c_int, _ctx_p, C.c_uint32, C.c_double, _dp, _dp)
_sig("nlc_batch_submit", C.c_int, _ctx_p, C.c_uint32, C.c_double)
_sig("nlc_batch_collect", C.c_int, _ctx_p, _dp, _dp)
_sig("nlc_set_batch_mode", C.c_int, _ctx_p, C.c_int)
BATCH_SLOTS = 4  # NLC_BATCH_SLOTS
_sig("nlc_autotune", C.c_int, _ctx_p, C.c_uint32, C.c_char_p, C.c_size_t)
_sig("nlc_ic_ragged_lengths", C.c_int, C.c_uint32, _u32p, C.c_uint32, C.c_uint64, _u32p)
_sig("nlc_ic_ragged_plane_waves", C.c_int, _u32p, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
     C.c_uint64, _dp)


class NlcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{lib.nlc_strerror(code).decode()} ({code}): {msg}")
        self.code = code


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _check(rc: int, ctx=None):
    if rc != NLC_OK:
        raise NlcError(rc, lib.nlc_last_error(ctx).decode(errors="replace"))


# ---- host helpers -----------------------------------------------------------

def yoshida() -> tuple[np.ndarray, np.ndarray]:
    c, d = np.zeros(8), np.zeros(8)
    _check(lib.nlc_yoshida(_ptr(c), _ptr(d)))
    return c, d


def dispersion(model: int, n: int, m: float = 0.0) -> np.ndarray:
    w = np.zeros(n)
    _check(lib.nlc_dispersion(model, n, m, _ptr(w)))
    return w


def eigensystem(mp2: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(omegas, V) with V[:, j] the j-th eigenvector (stored column-major)."""
    mp2 = np.ascontiguousarray(mp2, dtype=np.float64)
    n = mp2.size
    om = np.zeros(n)
    vt = np.zeros((n, n))  # row j = eigenvector j (column-major storage of V)
    _check(lib.nlc_eigensystem(n, _ptr(mp2), _ptr(om), _ptr(vt)))
    return om, vt.T.copy()


def ic_plane_waves(n: int, copies: int, m: float, k_lo: int, k_hi: int, e_density: float, seed: int) -> np.ndarray:
    s = np.zeros((copies, n, 2))
    _check(lib.nlc_ic_plane_waves(n, copies, m, k_lo, k_hi, e_density, seed, _ptr(s)))
    return s


def ic_dnls(n: int, copies: int, kmax: int, norm_density: float, seed: int) -> np.ndarray:
    s = np.zeros((copies, n, 2))
    _check(lib.nlc_ic_dnls(n, copies, kmax, norm_density, seed, _ptr(s)))
    return s


def ic_eigenmodes(n: int, copies: int, omegas: np.ndarray, V: np.ndarray, modes: int, e_density: float,
                  seed: int) -> np.ndarray:
    s = np.zeros((copies, n, 2))
    vt = np.ascontiguousarray(V.T)  # column-major V
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    _check(lib.nlc_ic_eigenmodes(n, copies, _ptr(om), _ptr(vt), modes, e_density, seed, _ptr(s)))
    return s


def ic_masses(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    m = np.zeros(n)
    _check(lib.nlc_ic_masses(n, lo, hi, seed, _ptr(m)))
    return m


def ic_ragged_lengths(nchains: int, choices, seed: int) -> np.ndarray:
    ch = np.ascontiguousarray(choices, dtype=np.uint32)
    out = np.zeros(nchains, dtype=np.uint32)
    _check(lib.nlc_ic_ragged_lengths(nchains, ch.ctypes.data_as(_u32p), ch.size, seed, out.ctypes.data_as(_u32p)))
    return out


def ic_ragged_plane_waves(lengths, m: float, mode_fraction: float, e_lo: float, e_hi: float, seed: int) -> np.ndarray:
    """Concatenated chains, float64[sum(lengths), 2]."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint32)
    s = np.zeros((int(ln.sum()), 2))
    _check(lib.nlc_ic_ragged_plane_waves(ln.ctypes.data_as(_u32p), ln.size, m, mode_fraction, e_lo, e_hi, seed,
                                         _ptr(s)))
    return s


# ---- ensembles on the GPU -----------------------------------------------------

class Ensemble:
    """One nlc_ctx: an ensemble of `copies` chains of length `n` resident on the GPU."""

    def __init__(self, model: int, n: int, copies: int, dt: float, *, m: float = 0.0, alpha: float = 0.0,
                 beta: float = 0.0, masses: np.ndarray | None = None, precision: int = FP64,
                 kernel: int = KERNEL_AUTO, lanes_per_chain: int = 0, sites_per_lane: int = 0,
                 projection: int = 0):
        p = Params()
        _check(lib.nlc_params_init(C.byref(p)))
        p.model, p.precision, p.kernel = model, precision, kernel
        p.chain_length, p.copies, p.dt, p.m, p.alpha, p.beta = n, copies, dt, m, alpha, beta
        p.lanes_per_chain, p.sites_per_lane, p.projection = lanes_per_chain, sites_per_lane, projection
        self._masses = None
        if masses is not None:
            self._masses = np.ascontiguousarray(masses, dtype=np.float64)
            p.masses = _ptr(self._masses)
        self.model, self.n, self.copies = model, n, copies
        self._ctx = _ctx_p()
        _check(lib.nlc_create(C.byref(self._ctx), C.byref(p)))
        self.lengths = None
        self._shape = (copies, n, 2)
        self._nlin, self._nent = n, 2

    @classmethod
    def ragged(cls, model: int, lengths, dt: float, *, m: float = 0.0, alpha: float = 0.0, beta: float = 0.0,
               precision: int = FP64):
        """A ragged ensemble: chains of the given lengths in one context."""
        return cls.mixed([dict(model=model, m=m, alpha=alpha, beta=beta, precision=precision)], lengths, None, dt)

    @classmethod
    def mixed(cls, sets, lengths, chain_set, dt: float):
        """A mixed ensemble: sets = [dict(model=..., m=..., alpha=..., beta=...)]
        (1 to 3), chain i of length lengths[i] following sets[chain_set[i]];
        chain_set None = a ragged ensemble of sets[0] (nlc_create_ragged)."""
        self = cls.__new__(cls)
        ps = (Params * len(sets))()
        for p, kw in zip(ps, sets):
            _check(lib.nlc_params_init(C.byref(p)))
            p.model, p.precision, p.dt = kw["model"], kw.get("precision", FP64), dt
            p.m, p.alpha, p.beta = kw.get("m", 0.0), kw.get("alpha", 0.0), kw.get("beta", 0.0)
        self._masses = None
        ln = np.ascontiguousarray(lengths, dtype=np.uint32)
        self._ctx = _ctx_p()
        if chain_set is None:
            _check(lib.nlc_create_ragged(C.byref(self._ctx), ps, ln.ctypes.data_as(_u32p), ln.size))
        else:
            cs = np.ascontiguousarray(chain_set, dtype=np.uint32)
            assert cs.size == ln.size
            _check(lib.nlc_create_mixed(C.byref(self._ctx), ps, len(sets), ln.ctypes.data_as(_u32p),
                                        cs.ctypes.data_as(_u32p), ln.size))
        self.model, self.lengths = sets[0]["model"], ln
        self.n, self.copies = int(ln.max()), int(ln.size)
        self._shape = (int(ln.sum()), 2)
        nc = C.c_uint32()
        _check(lib.nlc_ragged_classes(self._ctx, C.byref(nc), None, None, 0), self._ctx)
        cl, ct = np.zeros(nc.value, np.uint32), np.zeros(nc.value, np.uint32)
        _check(lib.nlc_ragged_classes(self._ctx, C.byref(nc), cl.ctypes.data_as(_u32p), ct.ctypes.data_as(_u32p),
                                      nc.value), self._ctx)
        self.class_lengths, self.class_counts = cl, ct
        cset = np.zeros(nc.value, np.uint32)
        _check(lib.nlc_ragged_class_sets(self._ctx, cset.ctypes.data_as(_u32p), nc.value), self._ctx)
        self.class_sets = cset
        self._nlin, self._nent = int(cl.sum()), 2 * nc.value
        return self

    def close(self):
        if self._ctx:
            lib.nlc_destroy(self._ctx)
            self._ctx = _ctx_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._ctx

    def upload(self, state: np.ndarray) -> None:
        state = np.ascontiguousarray(state, dtype=np.float64)
        assert state.shape == self._shape, state.shape
        _check(lib.nlc_upload(self._ctx, state.ctypes.data_as(C.c_void_p)), self._ctx)

    def download(self) -> np.ndarray:
        s = np.empty(self._shape)
        _check(lib.nlc_download(self._ctx, s.ctypes.data_as(C.c_void_p)), self._ctx)
        return s

    def move(self, steps: int) -> None:
        _check(lib.nlc_move(self._ctx, steps), self._ctx)

    def sync(self) -> None:
        _check(lib.nlc_sync(self._ctx), self._ctx)

    def set_entropy_mask(self, mask: np.ndarray | None, a0is0: bool) -> None:
        if mask is None:
            _check(lib.nlc_set_entropy_mask(self._ctx, None, int(a0is0)), self._ctx)
        else:
            mk = np.ascontiguousarray(mask, dtype=np.uint8)
            assert mk.size == self.n
            _check(lib.nlc_set_entropy_mask(self._ctx, mk.ctypes.data_as(C.c_void_p), int(a0is0)), self._ctx)

    def linenergies(self, norm: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        e, s = np.zeros(self._nlin), np.zeros(self._nent)
        _check(lib.nlc_linenergies(self._ctx, norm, _ptr(e), _ptr(s)), self._ctx)
        return e, s

    def batch(self, steps: int, norm: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
        """One dump interval (move + energies + entropies) from a cached hipGraph."""
        e, s = np.zeros(self._nlin), np.zeros(self._nent)
        _check(lib.nlc_batch(self._ctx, steps, norm, _ptr(e), _ptr(s)), self._ctx)
        return e, s

    def batch_submit(self, steps: int, norm: float = 0.0) -> None:
        """Queue one batch without waiting (at most BATCH_SLOTS queued)."""
        _check(lib.nlc_batch_submit(self._ctx, steps, norm), self._ctx)

    def batch_collect(self) -> tuple[np.ndarray, np.ndarray]:
        """Wait for the oldest queued batch; its energies and entropies."""
        e, s = np.zeros(self._nlin), np.zeros(self._nent)
        _check(lib.nlc_batch_collect(self._ctx, _ptr(e), _ptr(s)), self._ctx)
        return e, s

    def set_batch_mode(self, direct: bool) -> None:
        """Queued batches as graph replays (False, the default) or direct launches."""
        _check(lib.nlc_set_batch_mode(self._ctx, int(bool(direct))), self._ctx)

    def autotune(self, trial_steps: int = 20) -> str:
        buf = C.create_string_buffer(1024)
        _check(lib.nlc_autotune(self._ctx, trial_steps, buf, 1024), self._ctx)
        return buf.value.decode()

    def linenergies_raw(self) -> np.ndarray:
        e = np.zeros(self._nlin)
        _check(lib.nlc_linenergies_raw(self._ctx, _ptr(e)), self._ctx)
        return e

    def chain_energies(self) -> tuple[np.ndarray, np.ndarray]:
        e, nr = np.zeros(self.copies), np.zeros(self.copies)
        _check(lib.nlc_chain_energies(self._ctx, _ptr(e), _ptr(nr)), self._ctx)
        return e, nr

    def stream(self) -> int:
        return lib.nlc_stream(self._ctx) or 0

    def describe(self) -> str:
        buf = C.create_string_buffer(256)
        _check(lib.nlc_describe(self._ctx, buf, 256), self._ctx)
        return buf.value.decode()


def declared_symbols(header: Path | None = None) -> list[str]:
    """Function names declared in include/nlc.h."""
    import re
    header = header or (_PKG.parent / "include" / "nlc.h")
    text = header.read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(nlc_\w+)\s*\(", text, flags=re.M)))
