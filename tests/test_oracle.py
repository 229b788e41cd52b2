"""Pins the CPU oracle (oracle/oracle.c) with checks independent of its code:
published constants, closed forms, invariants, numpy/LAPACK fixtures and
brute force.  The reference ships no golden vectors or known-answer tests for
this path (its only validation is manual comparison of dumps), so every pin is
built here.
"""
import numpy as np
import pytest

from oracle import oracle as O

RNG = np.random.default_rng(12345)


# ---------------------------------------------------------------- constants

def test_yoshida_weights_from_published_solution_a():
    # Yoshida, Phys. Lett. A 150 (1990) 262, 6th order, solution A:
    w1, w2, w3 = -1.17767998417887, 0.235573213359357, 0.784513610477560
    w0 = 1 - 2 * (w1 + w2 + w3)
    c, d = O.yoshida()
    np.testing.assert_allclose(d[:7], [w3, w2, w1, w0, w1, w2, w3], rtol=0, atol=2e-15)
    assert d[7] == 0
    np.testing.assert_allclose(c[:4], [w3 / 2, (w3 + w2) / 2, (w2 + w1) / 2, (w1 + w0) / 2], rtol=0, atol=2e-15)
    assert list(c) == list(c[::-1])          # palindromic drift weights
    assert list(d[:7]) == list(d[:7][::-1])  # palindromic kick weights
    assert abs(c.sum() - 1) < 1e-14 and abs(d.sum() - 1) < 1e-14


# ---------------------------------------------------------- step: closed form

def yoshida_matrix(omega: float, dt: float) -> np.ndarray:
    """One Yoshida-6 step of the oscillator q'' = -omega^2 q as a 2x2 matrix."""
    c, d = O.yoshida()
    M = np.eye(2)
    for k in range(8):
        M = np.array([[1, c[k] * dt], [0, 1]]) @ M
        if k < 7:
            M = np.array([[1, 0], [-d[k] * dt * omega ** 2, 1]]) @ M
    return M


@pytest.mark.parametrize("model,n,m,steps,dt", [(O.DNKG, 16, 0.7, 40, 0.1), (O.FPUT, 12, 0.0, 25, 0.2),
                                                (O.DNKG, 64, 1.0, 100, 0.05)])
def test_move_linear_chain_matches_closed_form(model, n, m, steps, dt):
    """beta = alpha = 0: every Fourier mode is a harmonic oscillator; the
    Yoshida step is then an exact 2x2 product of shears, so n steps are
    pinned by a matrix power (up to rounding)."""
    s0 = RNG.normal(size=(3, n, 2))
    got = O.move_real(model, s0, steps, dt, m=m, alpha=0.0, beta=0.0)
    omega = np.sqrt(m + (2 * np.sin(np.pi * np.arange(n) / n)) ** 2)
    Q, P = np.fft.fft(s0[..., 0], axis=1), np.fft.fft(s0[..., 1], axis=1)
    for k in range(n):
        Mk = np.linalg.matrix_power(yoshida_matrix(omega[k], dt), steps)
        Q[:, k], P[:, k] = Mk[0, 0] * Q[:, k] + Mk[0, 1] * P[:, k], Mk[1, 0] * Q[:, k] + Mk[1, 1] * P[:, k]
    want = np.stack([np.fft.ifft(Q, axis=1).real, np.fft.ifft(P, axis=1).real], axis=-1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


# -------------------------------------------------------- forces vs -dH/dq

@pytest.mark.parametrize("model,kw", [(O.DNKG, dict(m=0.7, beta=1.3)), (O.FPUT, dict(alpha=0.4, beta=1.1)),
                                      (O.TODA, dict(alpha=0.6)), (O.DDNKG, dict(beta=0.9))])
def test_force_is_minus_energy_gradient(model, kw):
    n = 9
    q = RNG.normal(size=n) * 0.5
    masses = RNG.uniform(0.5, 1.5, size=n) if model == O.DDNKG else None
    mp2 = None if masses is None else masses + 2
    F = O.forces(model, q, mp2=mp2, **kw)
    h = 1e-5
    grad = np.zeros(n)
    for i in range(n):
        st = np.zeros((2, n, 2))
        st[0, :, 0] = q
        st[1, :, 0] = q
        st[0, i, 0] += h
        st[1, i, 0] -= h
        E = O.chain_energy(model, st, masses=masses, **kw)
        grad[i] = (E[0] - E[1]) / (2 * h)
    np.testing.assert_allclose(F, -grad, rtol=0, atol=1e-8 * max(1, np.abs(F).max()))


def test_chain_energy_hand_values():
    # README.md:15-19 Hamiltonians on a 3-site chain q = (0, h, 3h), p = (1, 0, -1)
    h = 0.3
    st = np.array([[[0, 1], [h, 0], [3 * h, -1]]], dtype=float)
    d = np.array([h, 2 * h, -3 * h])  # q_{x+1} - q_x, cyclic
    kin = 1.0
    q = st[0, :, 0]
    m, al, be = 0.5, 0.25, 0.75
    assert O.chain_energy(O.DNKG, st, m=m, beta=be)[0] == pytest.approx(
        kin + (d ** 2).sum() / 2 + m * (q ** 2).sum() / 2 + be * (q ** 4).sum() / 4, rel=1e-14)
    assert O.chain_energy(O.FPUT, st, alpha=al, beta=be)[0] == pytest.approx(
        kin + (d ** 2).sum() / 2 + al * (d ** 3).sum() / 3 + be * (d ** 4).sum() / 4, rel=1e-14)
    assert O.chain_energy(O.TODA, st, alpha=al)[0] == pytest.approx(
        kin + ((np.exp(2 * al * d) - 2 * al * d - 1) / (4 * al ** 2)).sum(), rel=1e-13)


def test_chain_energy_hand_values_dnls_ddnkg():
    """The DNLS and dDNKG branches, README.md:15-19 / nlchains.m:40-48, on
    3-site chains whose energies are exact binary fractions.
    DNLS: psi = (1, i, -1 + 2i), H = sum |psi_{x+1} - psi_x|^2 + beta/2 |psi_x|^4:
    gradient terms 2 + 2 + 8 = 12, quartic sum 1 + 1 + 25 = 27, beta = 0.5 gives
    12 + 6.75 = 18.75 (beta/4 instead of beta/2 would give 15.375).
    dDNKG: q = (0, h, 3h), p = (1, 0, -1), masses m_x = (0.5, 1.5, 1.0), h = 0.5,
    beta = 2: kinetic 1, bonds (1 + 4 + 9) h^2/2 = 1.75, masses
    (1.5 h^2 + 9 h^2)/2 = 1.3125, quartic beta (h^4 + 81 h^4)/4 = 2.5625:
    6.625 (the masses + 2 of the force in place of m_x would give 8.875)."""
    psi = np.array([[[1.0, 0.0], [0.0, 1.0], [-1.0, 2.0]]])
    assert O.chain_energy(O.DNLS, psi, beta=0.5)[0] == 18.75
    assert O.chain_energy(O.DNLS, psi, beta=0.0)[0] == 12.0
    h = 0.5
    st = np.array([[[0.0, 1.0], [h, 0.0], [3 * h, -1.0]]])
    masses = np.array([0.5, 1.5, 1.0])
    assert O.chain_energy(O.DDNKG, st, beta=2.0, masses=masses)[0] == 6.625
    assert O.chain_energy(O.DDNKG, st, beta=0.0, masses=masses)[0] == 6.625 - 2.5625


def test_move_matches_composition_of_pinned_force():
    """The kick loop of orc_move_real (its own expressions) against a numpy
    Yoshida composition that uses the FD-pinned orc_forces."""
    c, d = O.yoshida()
    n, dt, steps = 10, 0.07, 3
    for model, kw in [(O.DNKG, dict(m=0.7, beta=1.3)), (O.FPUT, dict(alpha=0.4, beta=1.1)),
                      (O.TODA, dict(alpha=0.6)), (O.DDNKG, dict(beta=0.9))]:
        s0 = RNG.normal(size=(1, n, 2)) * 0.4
        mp2 = RNG.uniform(2.5, 3.5, size=n) if model == O.DDNKG else None
        got = O.move_real(model, s0, steps, dt, mp2=mp2, **kw)
        q, p = s0[0, :, 0].copy(), s0[0, :, 1].copy()
        for _ in range(steps):
            for k in range(8):
                q = q + (c[k] * dt) * p
                if k < 7:
                    p = p + (d[k] * dt) * O.forces(model, q, mp2=mp2, **kw)
        np.testing.assert_allclose(got[0, :, 0], q, rtol=0, atol=1e-13)
        np.testing.assert_allclose(got[0, :, 1], p, rtol=0, atol=1e-13)


def test_energy_error_scales_as_dt6():
    """Truncation-dominated energy error: halving dt divides it by ~2^6
    (the reference's epsilon(dt) study, examples.nb:2576-2582)."""
    n, T = 32, 20.0
    s0 = RNG.normal(size=(8, n, 2)) * 0.5
    errs = []
    for dt in (0.2, 0.1):
        E0 = O.chain_energy(O.FPUT, s0, alpha=0.1, beta=0.1)
        s = O.move_real(O.FPUT, s0, int(round(T / dt)), dt, alpha=0.1, beta=0.1)
        errs.append(np.mean(np.abs(O.chain_energy(O.FPUT, s, alpha=0.1, beta=0.1) - E0) / E0))
    assert 20 < errs[0] / errs[1] < 200, errs


# ------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 64, 100, 1024])
def test_fft_matches_numpy(n):
    z = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    np.testing.assert_allclose(O.fft(z, -1), np.fft.fft(z), rtol=0, atol=1e-13 * np.abs(z).sum())
    np.testing.assert_allclose(O.fft(z, +1), n * np.fft.ifft(z), rtol=0, atol=1e-13 * np.abs(z).sum())


def test_fft_against_golden_fixture():
    g = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "fixtures.npz")
    for key in [k for k in g.files if k.startswith("fft_in_")]:
        n = key.split("_")[-1]
        np.testing.assert_allclose(O.fft(g[key]), g[f"fft_out_{n}"], rtol=0, atol=1e-12 * np.abs(g[key]).sum())


def test_fft_brute_force_dft():
    n = 16
    z = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    x = np.arange(n)
    want = np.array([(z * np.exp(-2j * np.pi * k * x / n)).sum() for k in range(n)])
    np.testing.assert_allclose(O.fft(z), want, rtol=0, atol=1e-13)


# --------------------------------------------------------- linear energies

def harmonic_energy(state, m, masses=None):
    q, p = state[..., 0], state[..., 1]
    d = np.roll(q, -1, axis=1) - q
    mass = m if masses is None else masses
    return (p ** 2 / 2 + d ** 2 / 2 + mass * q ** 2 / 2).sum(axis=1)


def test_linenergies_parseval():
    n, m = 32, 0.8
    s = RNG.normal(size=(5, n, 2))
    e = O.linenergies_real(s, O.dispersion_real(n, m))
    np.testing.assert_allclose(e.sum(), 2 * n * harmonic_energy(s, m).sum(), rtol=1e-13)
    assert (e >= -1e-12).all()


def test_linenergies_running_wave_is_one_hot():
    from nlchains_amd import ic_plane_waves  # product generator: only supplies the input
    n, m, k0, dens = 64, 1.0, 5, 0.3
    s = ic_plane_waves(n, 4, m, k0, k0 + 1, dens, 7)
    e = O.calc_linenergies(O.linenergies_real(s, O.dispersion_real(n, m)), 0.5 / 4 / n, False)
    assert e[k0] == pytest.approx(dens * n, rel=1e-12)
    assert np.abs(np.delete(e, k0)).max() < 1e-12 * e[k0]


def test_dispersion():
    n = 16
    k = np.arange(n)
    np.testing.assert_allclose(O.dispersion_real(n, 0.5), np.sqrt(0.5 + 4 * np.sin(np.pi * k / n) ** 2), rtol=1e-15)
    np.testing.assert_allclose(O.dispersion_dnls(n), 4 * np.sin(np.pi * k / n) ** 2, rtol=1e-15, atol=1e-15)


# ---------------------------------------------------------------- entropies

def test_entropy_equipartition_is_zero():
    wt, inf = O.entropies(np.full(17, 0.37), a0is0=False)
    assert abs(wt) < 1e-13 and abs(inf) < 1e-13
    e = np.full(17, 2.0)
    e[0] = 123.0  # excluded when a0is0
    wt, inf = O.entropies(e, a0is0=True)
    assert abs(wt) < 1e-13 and abs(inf) < 1e-13


def test_entropy_two_modes_closed_form():
    a, b = 3.0, 1.0  # f = (1.5, 0.5)
    wt, inf = O.entropies(np.array([a, b]), a0is0=False)
    assert wt == pytest.approx(-np.log(1.5) - np.log(0.5), rel=1e-14)
    assert inf == pytest.approx(1.5 * np.log(1.5) + 0.5 * np.log(0.5), rel=1e-14)


def test_entropy_properties():
    e = RNG.uniform(0.1, 2, size=50)
    wt, inf = O.entropies(e, a0is0=False)
    assert wt > 0 and inf > 0                     # AM-GM and Jensen
    wt2, inf2 = O.entropies(7.5 * e, a0is0=False)  # scale invariance
    assert wt2 == pytest.approx(wt, rel=1e-12) and inf2 == pytest.approx(inf, rel=1e-12)
    # masked subset equals the entropy of the subset alone
    mask = np.zeros(50, np.uint8)
    mask[[3, 7, 8, 20]] = 1
    idx = O.entropy_modes(50, mask, False)
    assert idx == [3, 7, 8, 20]
    assert O.entropies(e, False, idx) == pytest.approx(O.entropies(e[idx], False), rel=1e-14)


def test_entropy_mask_rules():
    mask = np.zeros(8, np.uint8)
    mask[[0, 5]] = 1
    assert O.entropy_modes(8, mask, False) == [0, 5]
    assert O.entropy_modes(8, mask, True) is None   # only one mode left after dropping 0
    mask[6] = 1
    assert O.entropy_modes(8, mask, True) == [5, 6]
    assert O.entropy_modes(8, np.eye(1, 8, 3, dtype=np.uint8)[0], False) is None


def test_calc_linenergies():
    e = np.arange(1.0, 6.0)
    np.testing.assert_array_equal(O.calc_linenergies(e, 0.5, True), [0, 1, 1.5, 2, 2.5])
    np.testing.assert_array_equal(O.calc_linenergies(e, 2.0, False), 2 * e)


# ------------------------------------------------------------ eigensystem

def test_eigensystem_matches_lapack_fixture():
    g = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "fixtures.npz")
    om, V, sweeps = O.eigensystem(g["eig_mp2"])
    assert sweeps > 0
    np.testing.assert_allclose(om, g["eig_omega"], rtol=1e-12)
    # eigenvectors up to sign: compare projectors onto each (non-degenerate) mode
    np.testing.assert_allclose(np.abs((V * g["eig_V"]).sum(axis=0)), 1, atol=1e-10)


def test_eigensystem_invariants_and_uniform_closed_form():
    n = 24
    mp2 = RNG.uniform(2.5, 3.5, size=n)
    om, V, _ = O.eigensystem(mp2)
    A = np.diag(mp2) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    np.testing.assert_allclose(A @ V, V * om ** 2, atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    assert (np.diff(om) >= 0).all()
    om_u, _, _ = O.eigensystem(np.full(n, 3.0))
    want = np.sort(np.sqrt(3.0 - 2 * np.cos(2 * np.pi * np.arange(n) / n)))
    np.testing.assert_allclose(om_u, want, rtol=1e-13)


def test_ddnkg_energies_harmonic_identity():
    n = 20
    masses = RNG.uniform(0.5, 1.5, size=n)
    om, V, _ = O.eigensystem(masses + 2)
    s = RNG.normal(size=(4, n, 2))
    e = O.linenergies_ddnkg(s, V, om)
    np.testing.assert_allclose(e.sum(), 2 * harmonic_energy(s, 0, masses).sum(), rtol=1e-12)
    # uniform masses: same spectrum as the plane-wave DNKG, summed over degenerate pairs
    om_u, V_u, _ = O.eigensystem(np.full(n, 2.7))
    e_pw = O.linenergies_real(s, O.dispersion_real(n, 0.7)) / n   # plane-wave norm carries an extra 1/n
    np.testing.assert_allclose(np.sort(np.round(om_u, 12)), np.sort(np.round(O.dispersion_real(n, 0.7), 12)))
    np.testing.assert_allclose(O.linenergies_ddnkg(s, V_u, om_u).sum(), e_pw.sum(), rtol=1e-12)


# -------------------------------------------------------------------- DNLS

def test_dnls_norm_conserved():
    n = 64
    s = RNG.normal(size=(3, n, 2)) * 0.7
    out = O.move_dnls(s, 50, 0.05, 1.3)
    np.testing.assert_allclose(O.dnls_norm(out), O.dnls_norm(s), rtol=1e-12)


def test_dnls_linear_closed_form():
    n, dt, steps = 32, 0.05, 40
    s = RNG.normal(size=(2, n, 2))
    out = O.move_dnls(s, steps, dt, 0.0)
    _, d = O.yoshida()
    w = O.dispersion_dnls(n)
    psi0 = np.fft.fft(s[..., 0] + 1j * s[..., 1], axis=1)
    want = np.fft.ifft(psi0 * np.exp(-1j * w * dt * steps * d[:7].sum()), axis=1)
    np.testing.assert_allclose(out[..., 0] + 1j * out[..., 1], want, atol=1e-12)


def test_dnls_plane_wave_exact_solution():
    """A single plane wave has uniform |psi|, so the nonlinear and linear flows
    commute and the split step is exact: psi = a e^{i(2 pi k x/n - (w_k + beta|a|^2) t)}."""
    n, k0, a, beta, dt, steps = 16, 3, 0.8, 1.7, 0.02, 30
    c, d = O.yoshida()
    x = np.arange(n)
    psi = a * np.exp(2j * np.pi * k0 * x / n)
    s = np.stack([psi.real, psi.imag], axis=-1)[None]
    out = O.move_dnls(s, steps, dt, beta)
    wk = O.dispersion_dnls(n)[k0]
    phase = wk * dt * steps * d[:7].sum() + beta * a * a * dt * steps * c.sum()
    want = psi * np.exp(-1j * phase)
    np.testing.assert_allclose(out[0, :, 0] + 1j * out[0, :, 1], want, atol=1e-13)


def test_dnls_energies_definition():
    n = 16
    s = RNG.normal(size=(3, n, 2))
    e = O.linenergies_dnls(s)
    psi_k = np.fft.fft(s[..., 0] + 1j * s[..., 1], axis=1)
    np.testing.assert_allclose(e, O.dispersion_dnls(n) * (np.abs(psi_k) ** 2).sum(axis=0), rtol=1e-12, atol=1e-12)


def test_dnls_tables():
    n, dt = 8, 0.1
    t = O.dnls_tables(n, dt)
    _, d = O.yoshida()
    want = np.exp(-1j * dt * d[:7, None] * O.dispersion_dnls(n)[None, :]) / n
    np.testing.assert_allclose(t, want, atol=1e-16)
