"""nlc_linenergies_noalloc: the allocation-free energy pass (static device
scratch, no hipMalloc) against nlc_linenergies of a context holding the same
state, through the C ABI: two calls in a row, the second with a smaller input,
and the size limits.  Named to be collected last: it must pass alone in a
fresh process and after every other GPU test."""
import ctypes as C

import numpy as np
import pytest

from parity import vec_rel_err

pytestmark = pytest.mark.gpu

MAX_CHAIN = 4096  # NLC_NOALLOC_MAX_CHAIN


class Hip:
    def __init__(self):
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h = h

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def to_host(self, dptr, count):
        out = np.empty(count)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), dptr, out.nbytes, 2) == 0  # DeviceToHost
        return out


def noalloc(N):
    f = N.lib.nlc_linenergies_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_int,
                  C.c_void_p, C.c_void_p]
    return f


def case(N, model, n, copies, seed):
    if model == N.DNLS:
        s = N.ic_dnls(n, copies, max(1, n // 16), 1.0, seed)
        kw = dict(beta=1.0)
    else:
        m = 1.0 if model == N.DNKG else 0.0
        s = N.ic_plane_waves(n, copies, m, 0 if m else 1, max(2, n // 8), 0.5, seed)
        kw = dict(m=m, beta=1.0) if model == N.DNKG else dict(alpha=1.0) if model == N.TODA else dict(beta=1.0)
    # every mode carries energy: finite entropies
    s = s + 0.01 * np.random.default_rng(seed).normal(size=s.shape)
    return s, kw


@pytest.mark.parametrize("first,second", [
    ((0, 4096, 600), (0, 1024, 37)),   # DNKG, then a smaller input (fewer, shorter chains)
    ((4, 4096, 300), (4, 256, 5)),     # DNLS
    ((1, 2048, 128), (2, 64, 1000)),   # FPUT, then Toda on 64-site chains (small-chain path)
])
def test_noalloc_matches_context_path(gpu, first, second):
    N = gpu
    hip = Hip()
    f = noalloc(N)
    out = hip.malloc(8 * (MAX_CHAIN + 2))
    try:
        for model, n, copies in (first, second):
            s0, kw = case(N, model, n, copies, 3 + n)
            with N.Ensemble(model, n, copies, 0.05, **kw) as ens:
                ens.upload(s0)
                want_e, want_s = ens.linenergies()
                dptr, nbytes = C.c_void_p(), C.c_size_t()
                assert N.lib.nlc_device_state(ens.handle, C.byref(dptr), C.byref(nbytes)) == 0
                a0is0 = 1 if model == N.DNLS or kw.get("m", 0.0) == 0.0 else 0
                rc = f(model, dptr, n, copies, kw.get("m", 0.0), 0.0, a0is0, out, C.c_void_p(ens.stream()))
                assert rc == 0, N.lib.nlc_last_error(None)
                ens.sync()
                got = hip.to_host(out, n + 2)
            e_err = vec_rel_err(got[:n], want_e)
            s_err = np.abs(got[n:] / want_s - 1).max()
            print(f"[gpu-noalloc] model={model} n={n} copies={copies}: linenergies {e_err:.2e} entropies {s_err:.2e}",
                  flush=True)
            assert e_err <= 1e-12 and s_err <= 1e-12
    finally:
        hip.h.hipFree(out)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    N = gpu
    f = noalloc(N)
    hip = Hip()
    buf = hip.malloc(8 * 2 * 8192 * 2)
    try:
        for model, n, copies, code in ((0, 8192, 2, N.NLC_EARG), (0, 100, 2, N.NLC_EARG), (0, 4, 2, N.NLC_EARG),
                                       (0, 64, 0, N.NLC_EARG), (3, 64, 2, N.NLC_EUNSUPPORTED)):
            rc = f(model, buf, n, copies, 1.0, 0.0, 0, buf, None)
            assert rc == code, (model, n, copies, rc)
            assert b"nlc_linenergies_noalloc" in N.lib.nlc_last_error(None)
        assert f(0, None, 64, 2, 1.0, 0.0, 0, buf, None) == N.NLC_EARG
    finally:
        hip.h.hipFree(buf)
