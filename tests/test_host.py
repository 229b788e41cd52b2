"""CPU tests of libnlchains_amd: the C ABI loads and exports everything
include/nlc.h declares, the host helpers are right, and device entry points
fail cleanly (no CPU fallback) where there is no GPU."""
import ctypes
import os

import numpy as np
import pytest

import nlchains_amd as N
from oracle import oracle as O


def test_abi_exports_every_declared_symbol():
    declared = N.declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(N.lib, name), f"{name} declared in include/nlc.h but not exported"
    assert N.lib.nlc_abi_version() == 1


def test_library_links_hip_runtime_and_no_oracle():
    maps = open(f"/proc/{os.getpid()}/maps").read()
    assert "libnlchains_amd.so" in maps
    # the product library must not depend on the oracle
    import subprocess
    deps = subprocess.run(["ldd", str(N.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in deps and "libamdhip64" in deps


def test_strerror_and_params_init():
    assert N.lib.nlc_strerror(0) == b"success"
    assert N.lib.nlc_strerror(-1) == b"invalid argument"
    p = N.Params()
    assert N.lib.nlc_params_init(ctypes.byref(p)) == 0
    assert p.struct_size == ctypes.sizeof(N.Params) and p.model == N.DNKG and p.precision == N.FP64


@pytest.mark.parametrize("kw,code", [
    (dict(n=1), N.NLC_EARG), (dict(copies=0), N.NLC_EARG), (dict(dt=0.0), N.NLC_EARG),
    (dict(dt=-1.0), N.NLC_EARG), (dict(model=N.TODA, alpha=0.0), N.NLC_EARG),
    (dict(model=N.DDNKG), N.NLC_EARG), (dict(n=1 << 16, copies=(1 << 15) + 1), N.NLC_EARG),
    (dict(precision=7), N.NLC_EARG),
    # FP32C: DNLS, FPUT and Toda only
    (dict(model=N.DNKG, precision=N.FP32C), N.NLC_EUNSUPPORTED),
    (dict(model=N.DDNKG, masses=np.ones(64), precision=N.FP32C), N.NLC_EUNSUPPORTED),
])
def test_create_rejects_invalid_arguments(kw, code):
    """The argument rules of common/main.cpp:83-103 (checked before any device call)."""
    args = dict(model=N.DNKG, n=64, copies=4, dt=0.1, alpha=1.0)
    args.update(kw)
    model, n, copies, dt = args.pop("model"), args.pop("n"), args.pop("copies"), args.pop("dt")
    with pytest.raises(N.NlcError) as ei:
        N.Ensemble(model, n, copies, dt, **args)
    assert ei.value.code == code


def test_create_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd") and N.lib.nlc_abi_version() and _have_device():
        pytest.skip("a GPU is present")
    with pytest.raises(N.NlcError) as ei:
        N.Ensemble(N.DNKG, 64, 4, 0.1, m=1.0, beta=1.0)
    assert ei.value.code == N.NLC_EHIP and "no CPU fallback" in str(ei.value)


def _have_device():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_yoshida_and_dispersion_match_oracle():
    c, d = N.yoshida()
    oc, od = O.yoshida()
    assert (c == oc).all() and (d == od).all()
    for n in (2, 7, 64):
        np.testing.assert_array_equal(N.dispersion(N.DNKG, n, 0.3), O.dispersion_real(n, 0.3))
        np.testing.assert_array_equal(N.dispersion(N.FPUT, n, 0.3), O.dispersion_real(n, 0.0))
        np.testing.assert_array_equal(N.dispersion(N.DNLS, n), O.dispersion_dnls(n))


@pytest.mark.parametrize("n", [2, 3, 17, 64, 200, 511, 512])
def test_product_eigensystem_vs_lapack_and_oracle(n):
    rng = np.random.default_rng(n)
    mp2 = 2 + rng.uniform(0.5, 1.5, size=n)
    om, V = N.eigensystem(mp2)
    A = np.diag(mp2) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    lam = np.linalg.eigvalsh(A)
    np.testing.assert_allclose(om, np.sqrt(lam), rtol=1e-12)
    np.testing.assert_allclose(A @ V, V * om ** 2, atol=1e-11)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    if n <= 64:
        om_o, _, _ = O.eigensystem(mp2)
        np.testing.assert_allclose(om, om_o, rtol=1e-12)


def test_eigensystem_uniform_masses_degenerate_spectrum():
    n = 32
    om, V = N.eigensystem(np.full(n, 3.0))
    np.testing.assert_allclose(om, np.sort(np.sqrt(3 - 2 * np.cos(2 * np.pi * np.arange(n) / n))), rtol=1e-13)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    A = 3 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    np.testing.assert_allclose(A @ V, V * om ** 2, atol=1e-12)


def test_eigensystem_nearly_degenerate_and_graded_masses():
    """Masses that split the degenerate pairs of the uniform chain by 1e-9,
    and masses graded over three orders of magnitude: eigenvalues against
    LAPACK, eigen-equation and orthonormality."""
    n = 96
    for mp2 in (3.0 + 1e-9 * np.sin(np.arange(n)), 2 + np.geomspace(1e-2, 10.0, n)):
        om, V = N.eigensystem(mp2)
        A = np.diag(mp2) - np.eye(n, k=1) - np.eye(n, k=-1)
        A[0, -1] = A[-1, 0] = -1
        np.testing.assert_allclose(om ** 2, np.linalg.eigvalsh(A), rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(A @ V, V * om ** 2, atol=1e-11)
        np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)


def test_ic_plane_waves_energy_and_determinism():
    n, copies, m, dens = 128, 6, 1.0, 0.5
    s = N.ic_plane_waves(n, copies, m, 0, 16, dens, 2)
    assert s.shape == (copies, n, 2)
    np.testing.assert_array_equal(s, N.ic_plane_waves(n, copies, m, 0, 16, dens, 2))
    assert not np.array_equal(s, N.ic_plane_waves(n, copies, m, 0, 16, dens, 3))
    # each chain carries harmonic energy dens*n, spread equally over modes 0..15
    e = O.calc_linenergies(O.linenergies_real(s, O.dispersion_real(n, m)), 0.5 / copies / n, False)
    np.testing.assert_allclose(e[:16], dens * n / 16, rtol=1e-10)
    assert np.abs(e[16:]).max() < 1e-10


def test_ic_dnls_norm_density():
    n, copies = 256, 3
    s = N.ic_dnls(n, copies, 16, 1.0, 3)
    np.testing.assert_allclose(O.dnls_norm(s) / n, 1.0, rtol=1e-12)
    psi_k = np.fft.fft(s[..., 0] + 1j * s[..., 1], axis=1)
    k = np.fft.fftfreq(n, 1 / n)
    assert np.abs(psi_k[:, np.abs(k) >= 16]).max() < 1e-9


def test_ic_eigenmodes_energy():
    n, copies, modes, dens = 48, 4, 6, 0.5
    masses = N.ic_masses(n, 0.5, 1.5, 5)
    assert masses.min() >= 0.5 and masses.max() < 1.5
    om, V = N.eigensystem(masses + 2)
    s = N.ic_eigenmodes(n, copies, om, V, modes, dens, 5)
    e = O.linenergies_ddnkg(s, V, om) * 0.5 / copies
    np.testing.assert_allclose(e[:modes], dens * n / modes, rtol=1e-10)
    assert np.abs(e[modes:]).max() < 1e-10


def test_ic_rejects_bad_arguments():
    with pytest.raises(N.NlcError):
        N.ic_plane_waves(64, 1, 1.0, 10, 5, 0.1, 1)
    with pytest.raises(N.NlcError):
        N.ic_dnls(16, 1, 9, 1.0, 1)
