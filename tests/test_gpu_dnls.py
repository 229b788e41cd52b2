"""GPU parity of the DNLS split-step integrator (FP64) against the oracle.

Bars: psi after 100 steps within 1e-10 relative (per chain: max-norm error over
the chain's max-norm, largest over all chains); normalised linear
energies and entropies within 1e-9; norm conserved to 1e-12 relative.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors, chain_rel_err, oracle_parallel, vec_rel_err

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-dnls] {msg}", flush=True)
    sys.stdout.flush()


def rel_err(got, want):
    """Largest per-chain relative max-norm error of two states."""
    return chain_rel_err(got, want)


def run(N, n, copies, steps, dt, beta, s0, spl=0):
    """spl: values per thread of the 4096-point register kernel (16 or 8; 0 = default)."""
    with N.Ensemble(N.DNLS, n, copies, dt, beta=beta, sites_per_lane=spl) as ens:
        ens.upload(s0)
        ens.move(steps)
        got = ens.download()
        desc = ens.describe()
    return got, desc


@pytest.mark.parametrize("n,copies,steps", [(4096, 6, 100), (2, 5, 100), (16, 5, 100), (64, 7, 100),
                                            (256, 5, 100), (1024, 3, 100), (2048, 3, 60), (8192, 3, 20),
                                            # 1024/n chains per LDS workgroup: several groups, the
                                            # last one partly filled
                                            (2, 600, 50), (8, 300, 50), (64, 100, 20), (512, 5, 30),
                                            # mixed radix, several chains per workgroup with left-over
                                            # threads: 25 and 125 threads per chain
                                            (100, 33, 20), (500, 5, 20), (6, 300, 30),
                                            # mixed-radix Stockham (factors 2, 3, 5, 7 only)
                                            (3, 5, 100), (5, 4, 100), (6, 5, 100), (7, 4, 100), (100, 5, 100),
                                            (1000, 3, 30), (1225, 2, 8), (3000, 2, 5), (6561, 1, 1),
                                            (7000, 1, 1), (8000, 1, 1),
                                            # Bluestein: convolution sizes 32, 256, 4096, 8192
                                            (11, 4, 100), (101, 5, 50), (1009, 2, 20), (4095, 2, 3),
                                            # two-level FFT in HBM: 4 x 4096, 8 x 4096, 16 x 4096, 16 x 8192
                                            (16384, 3, 10), (32768, 2, 6), (65536, 2, 3), (131072, 1, 3),
                                            # Bluestein through the two-level FFT (L = 16384, 16384, 65536)
                                            (5003, 2, 5), (6007, 2, 3), (20001, 1, 1)])
def test_dnls_move_parity(gpu, n, copies, steps):
    N = gpu
    dt, beta = 0.01, 1.0
    s0 = N.ic_dnls(n, copies, max(1, min(n // 16, 256)), 1.0, 3)
    banner(f"move n={n} copies={copies} steps={steps}")
    got, desc = run(N, n, copies, steps, dt, beta, s0)
    want = O.move_dnls(s0, steps, dt, beta)
    err = rel_err(got, want)
    banner(f"  {desc}: rel err {err:.3e}")
    assert np.isfinite(got).all() and err <= 1e-10, (desc, err)
    np.testing.assert_allclose(O.dnls_norm(got), O.dnls_norm(s0), rtol=1e-12)


@pytest.mark.parametrize("spl", [16, 8])
def test_dnls_4096_register_kernels(gpu, spl):
    """Both 4096-point register kernels (radix 16 x 256 threads, radix 8 x 512
    threads) against the oracle, and against each other: two FP64 roundings of
    the same transforms (the radix-16 kernel forms half of its twiddle powers
    by a three-term recurrence), measured 1.1e-12 apart after 100 steps."""
    N = gpu
    n, copies, steps, dt, beta = 4096, 5, 100, 0.01, 1.0
    s0 = N.ic_dnls(n, copies, 256, 1.0, 7)
    got, desc = run(N, n, copies, steps, dt, beta, s0, spl)
    banner(f"  {desc}")
    assert ("radix8" in desc) == (spl == 8)
    want = O.move_dnls(s0, steps, dt, beta)
    assert np.isfinite(got).all() and rel_err(got, want) <= 1e-10
    np.testing.assert_allclose(O.dnls_norm(got), O.dnls_norm(s0), rtol=1e-12)
    other, _ = run(N, n, copies, steps, dt, beta, s0, 24 - spl)
    assert rel_err(got, other) <= 2e-12


@pytest.mark.parametrize("spl", [16, 8])
def test_dnls_large_angles_use_library_sincos(gpu, spl):
    """beta large enough that |theta| > pi/4 at many sites: the library sincos
    branch of the nonlinear rotation is exercised."""
    N = gpu
    n, copies, dt, beta = 4096, 3, 0.02, 60.0
    s0 = N.ic_dnls(n, copies, 64, 1.0, 11)
    assert (beta * dt * 2 * 0.3922568 * (s0 ** 2).sum(-1)).max() > np.pi / 4
    # strongly nonlinear, so round-off grows fast: few steps suffice to test the branch
    got, _ = run(N, n, copies, 3, dt, beta, s0, spl)
    want = O.move_dnls(s0, 3, dt, beta)
    assert rel_err(got, want) <= 1e-12


@pytest.mark.parametrize("spl", [16, 8])
def test_dnls_batches_compose(gpu, spl):
    N = gpu
    s0 = N.ic_dnls(4096, 2, 256, 1.0, 5)
    a, _ = run(N, 4096, 2, 30, 0.01, 1.0, s0, spl)
    with N.Ensemble(N.DNLS, 4096, 2, 0.01, beta=1.0, sites_per_lane=spl) as ens:
        ens.upload(s0)
        ens.move(10)
        ens.move(20)
        b = ens.download()
    assert rel_err(a, b) < 1e-13


@pytest.mark.parametrize("n,copies", [(4096, 5), (256, 9), (64, 3), (100, 7), (8192, 2), (16384, 3), (131072, 2),
                                      (6007, 3), (33001, 1)])
def test_dnls_linenergies(gpu, n, copies):
    N = gpu
    s0 = N.ic_dnls(n, copies, max(1, min(n // 16, 256)), 1.0, 4)
    s0 = s0 + 0.01 * np.random.default_rng(3).normal(size=s0.shape)
    with N.Ensemble(N.DNLS, n, copies, 0.01, beta=1.0) as ens:
        ens.upload(s0)
        e_gpu, s_gpu = ens.linenergies()
        E, norms = ens.chain_energies()
    e = O.calc_linenergies(O.linenergies_dnls(s0), 1.0 / (n * copies), True)
    wt, inf = O.entropies(e, True)
    assert vec_rel_err(e_gpu, e) <= 1e-9
    assert abs(s_gpu[0] - wt) <= 1e-9 * abs(wt) and abs(s_gpu[1] - inf) <= 1e-9 * abs(inf)
    np.testing.assert_allclose(E, O.chain_energy(O.DNLS, s0, beta=1.0), rtol=1e-12)
    np.testing.assert_allclose(norms, O.dnls_norm(s0), rtol=1e-12)


@pytest.mark.parametrize("spl", [16, 8])
def test_dnls_full_size_norm_and_sampled_parity(gpu, spl):
    """4096 sites x 4096 copies: norm conservation (both substeps are unitary),
    and a sample of chains re-run on the oracle."""
    N = gpu
    n, copies, steps, dt = 4096, 4096, 50, 0.01
    banner(f"full-size DNLS 4096 x 4096, {spl} values per thread")
    s0 = N.ic_dnls(n, copies, 256, 1.0, 3)
    with N.Ensemble(N.DNLS, n, copies, dt, beta=1.0, sites_per_lane=spl) as ens:
        ens.upload(s0)
        E0, n0 = ens.chain_energies()
        ens.move(steps)
        got = ens.download()
        E1, n1 = ens.chain_energies()
    np.testing.assert_allclose(n1, n0, rtol=1e-12)
    pick = np.sort(np.random.default_rng(1).choice(copies, 64, replace=False))
    want = oracle_parallel(O.move_dnls, s0[pick], steps=steps, dt=dt, beta=1.0)
    err = chain_errors(got[pick], want)
    banner(f"  64 sampled chains: largest per-chain rel err {err.max():.3e}")
    assert err.max() <= 1e-10
    E1o = O.chain_energy(O.DNLS, want, beta=1.0)
    np.testing.assert_allclose(E1[pick], E1o, rtol=1e-11)
    banner(f"  norm drift {np.abs(n1 / n0 - 1).max():.2e}, energy error {np.abs(E1 / E0 - 1).mean():.2e}")


def test_dnls_autotune_and_unsupported_shapes(gpu):
    N = gpu
    s0 = N.ic_dnls(4096, 64, 256, 1.0, 2)
    with N.Ensemble(N.DNLS, 4096, 64, 0.01, beta=1.0) as ens:
        ens.upload(s0)
        report = ens.autotune(5)
        banner(f"autotune: {report} -> {ens.describe()}")
        assert "radix16x256" in report and "radix8x512" in report and "*" in report
        assert np.array_equal(ens.download(), s0)
    with pytest.raises(N.NlcError):
        N.Ensemble(N.DNLS, 4096, 4, 0.01, beta=1.0, sites_per_lane=4)
    with pytest.raises(N.NlcError):
        N.Ensemble(N.DNLS, 1024, 4, 0.01, beta=1.0, sites_per_lane=8)
    for n in (1, 65537, 262144):  # no kernel: not a power of two above 65536, or above 131072
        with pytest.raises(N.NlcError):
            N.Ensemble(N.DNLS, n, 4, 0.01, beta=1.0)


@pytest.mark.parametrize("n", [33001, 50000])
def test_dnls_bluestein_two_level_linear_closed_form(gpu, n):
    """beta = 0 (L = 131072 = 16 x 8192): every Fourier mode only turns its
    phase, psi_hat_k(t) = psi_hat_k(0) e^{-i omega_k t} (numpy's FFT as the
    reference); beta = 1: the norm of every chain is conserved."""
    N = gpu
    s0 = N.ic_dnls(n, 2, 128, 1.0, 6) + 1e-3 * np.random.default_rng(5).normal(size=(2, n, 2))
    steps, dt = 4, 0.01
    with N.Ensemble(N.DNLS, n, 2, dt, beta=0.0) as ens:
        banner(f"  {ens.describe()}")
        assert "bluestein_two_level" in ens.describe()
        ens.upload(s0)
        ens.move(steps)
        got = ens.download()
    psi0 = s0[..., 0] + 1j * s0[..., 1]
    om = O.dispersion_dnls(n)
    want = np.fft.ifft(np.fft.fft(psi0, axis=1) * np.exp(-1j * om * dt * steps), axis=1)
    g = got[..., 0] + 1j * got[..., 1]
    assert chain_rel_err(g, want) <= 1e-11
    with N.Ensemble(N.DNLS, n, 2, dt, beta=1.0) as ens:
        ens.upload(s0)
        _, n0 = ens.chain_energies()
        ens.move(steps)
        _, n1 = ens.chain_energies()
    np.testing.assert_allclose(n1, n0, rtol=1e-12)


@pytest.mark.parametrize("n", [101, 2048, 8192, 32768, 5003, 100, 5000])
def test_dnls_generic_kernels_describe_and_compose(gpu, n):
    """The Bluestein, LDS Stockham, two-level (8192 sites: two rows of 4096)
    and mixed-radix kernels: kernel names, and two batches of 10 steps equal
    one of 20 to rounding."""
    N = gpu
    s0 = N.ic_dnls(n, 3, max(1, min(n // 16, 256)), 1.0, 5)
    with N.Ensemble(N.DNLS, n, 3, 0.01, beta=1.0) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(10)
        ens.move(10)
        a = ens.download()
        ens.upload(s0)
        ens.move(20)
        b = ens.download()
    banner(f"  {desc}")
    assert ("lds_bluestein" in desc) == (n == 101) and ("stockham" in desc) == (n == 2048)
    assert ("two_level_fft" in desc) == (n in (8192, 32768)) and ("bluestein_two_level" in desc) == (n == 5003)
    assert ("mixed_radix" in desc) == (n in (100, 5000))
    assert rel_err(a, b) < 1e-12
