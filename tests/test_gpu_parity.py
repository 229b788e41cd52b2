"""GPU parity of the FP64 hot path against the CPU oracle, through the C ABI.

Bars (the plain-form acceptance of the build):
  * q, p after 100 steps within 1e-10 relative, chain by chain (the max-norm
    error of each chain over that chain's max-norm, largest over all chains);
  * normalised linear energies and both entropies within 1e-9 relative;
  * chain energies within 1e-12 relative.
Inputs are seeded synthetic ensembles (the reference ships no initial states).
"""
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors, chain_rel_err, oracle_parallel, vec_rel_err

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10
ENERGY_TOL = 1e-9


def banner(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


MODEL_KW = {
    "DNKG": dict(m=1.0, beta=1.0),
    "FPUT": dict(alpha=0.25, beta=1.0),
    "Toda": dict(alpha=1.0),
    "dDNKG": dict(beta=1.0),
}


def make_case(N, model_name, n, copies, seed, dens=0.5, kw=None, plane=False):
    """Seeded initial state; kw overrides the model parameters, plane=True
    gives dDNKG plane waves instead of eigenmodes (no eigensystem)."""
    model = N.MODELS[model_name]
    kw = dict(MODEL_KW[model_name] if kw is None else kw)
    masses = None
    if model == N.DDNKG and plane:
        masses = N.ic_masses(n, 0.5, 1.5, seed)
        s = N.ic_plane_waves(n, copies, 1.0, 0, max(1, n // 8), dens, seed)
    elif model == N.DDNKG:
        masses = N.ic_masses(n, 0.5, 1.5, seed)
        om, V = N.eigensystem(masses + 2)
        s = N.ic_eigenmodes(n, copies, om, V, max(1, n // 8), dens, seed)
    else:
        m = kw.get("m", 0.0)
        k_lo = 0 if m > 0 else 1
        s = N.ic_plane_waves(n, copies, m, k_lo, max(k_lo + 1, n // 8), dens, seed)
    return model, kw, masses, s


def oracle_move(model, kw, masses, s, steps, dt):
    mp2 = None if masses is None else masses + 2
    okw = {k: v for k, v in kw.items()}
    return O.move_real(model, s, steps, dt, mp2=mp2, threads=1, **okw)


def check_move(N, model_name, n, copies, steps=100, dt=0.05, seed=1, kw=None, dens=0.5, plane=False, expect=None,
               **shape):
    banner(f"move {model_name} n={n} copies={copies} steps={steps} kw={kw} shape={shape}")
    model, kw, masses, s0 = make_case(N, model_name, n, copies, seed, dens=dens, kw=kw, plane=plane)
    with N.Ensemble(model, n, copies, dt, masses=masses, **kw, **shape) as ens:
        ens.upload(s0)
        ens.move(steps)
        got = ens.download()
        desc = ens.describe()
    want = oracle_move(model, kw, masses, s0, steps, dt)
    err = chain_rel_err(got, want)
    banner(f"  {desc}: rel err {err:.3e}")
    assert np.isfinite(got).all()
    assert err <= STATE_TOL, (desc, err)
    if expect is not None:
        assert expect in desc, desc
    return got


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n,copies", [(64, 96), (1024, 40)])
def test_move_default_shape(gpu, model_name, n, copies):
    check_move(gpu, model_name, n, copies)


@pytest.mark.parametrize("n,L,S", [(64, 8, 8), (64, 16, 4), (64, 32, 2), (64, 64, 1), (64, 2, 32), (64, 4, 16),
                                   (32, 1, 32), (32, 4, 8), (16, 2, 8), (8, 1, 8), (2, 1, 2), (128, 64, 2),
                                   (2048, 64, 32), (512, 16, 32), (256, 8, 32)])
def test_move_wave_shapes(gpu, n, L, S):
    check_move(gpu, "DNKG", n, 37, steps=60, lanes_per_chain=L, sites_per_lane=S)
    check_move(gpu, "FPUT", n, 37, steps=60, lanes_per_chain=L, sites_per_lane=S)


@pytest.mark.parametrize("n,L,S", [(64, 4, 16), (1024, 64, 16), (1500, 64, 24), (64, 2, 32), (25, 1, 25),
                                   (30, 1, 30), (512, 16, 32), (500, 16, 32), (2048, 64, 32), (2000, 64, 32)])
def test_toda_exp_forms(gpu, n, L, S):
    """Toda on both exponential forms: up to 24 sites per lane the table form
    on v = (256 alpha/ln2) q, from 25 on the table-free form on the state
    scaled to octaves (v/32); 100 steps against the oracle.  Energy density
    0.25: at 0.5 the 2048-site plane waves of this seed start with bond
    stretches up to 3.4 and one chain goes non-finite within 100 steps of
    dt = 0.05 in the oracle too."""
    check_move(gpu, "Toda", n, 23, steps=100, dens=0.25, lanes_per_chain=L, sites_per_lane=S)


# Parameters that select each force form of the register-resident kernels;
# the scaled forms hold sigma q, sigma p with sigma = sqrt(beta) (cubic) or
# alpha (quadratic), != 1 here.  n = 12: thread per chain; 1000 / 1024: wave
# kernel with ragged / full lanes; 3000 / 4096: workgroup kernel.
FORM_CASES = [
    ("DNKG", dict(m=1.0, beta=0.3), "cubic_unit"),
    ("DNKG", dict(m=0.7, beta=0.0), "plain"),
    ("DNKG", dict(m=1.0, beta=-0.02), "plain"),
    ("FPUT", dict(alpha=0.0, beta=2.5), "cubic_unit"),
    ("FPUT", dict(alpha=0.3, beta=0.0), "quadratic_unit"),
    ("FPUT", dict(alpha=-0.4, beta=0.0), "quadratic_unit"),
    ("FPUT", dict(alpha=0.25, beta=0.7), "plain"),
    ("dDNKG", dict(beta=0.4), "cubic_unit"),
    ("dDNKG", dict(beta=0.0), "plain"),
]


@pytest.mark.parametrize("model_name,kw,form", FORM_CASES)
@pytest.mark.parametrize("n", [12, 1000, 1024, 3000, 4096])
def test_force_forms(gpu, model_name, kw, form, n):
    check_move(gpu, model_name, n, 6, steps=60, kw=kw, dens=0.1, plane=True, expect=f"force={form}")


@pytest.mark.parametrize("n", [65, 100, 1000, 1500, 2047])
@pytest.mark.parametrize("model_name", ["DNKG", "Toda", "dDNKG"])
def test_move_ragged_lanes(gpu, model_name, n):
    """n not a multiple of 64: the last lanes hold one site fewer."""
    check_move(gpu, model_name, n, 21, steps=50)


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n", [3, 64, 2500])
def test_move_split_kernel(gpu, model_name, n):
    """The HBM-streaming kernel (the reference's --split_kernel)."""
    check_move(gpu, model_name, n, 33, steps=40, kernel=gpu.KERNEL_SPLIT)


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n,steps", [(16385, 40), (20000, 36)])
def test_move_tiled_default(gpu, model_name, n, steps):
    """Chains longer than 16384 sites run on the temporally blocked kernel by
    default (launches of 18 steps, 9 for Toda: odd and even launch counts)."""
    check_move(gpu, model_name, n, 2, steps=steps, plane=True, expect="kernel=tiled")


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n", [300, 1000, 5000])
def test_move_tiled_forced(gpu, model_name, n):
    """Forced on shorter chains; at 300 and 1000 sites a segment (2048 sites,
    1024 for Toda) wraps around the ring more than once."""
    check_move(gpu, model_name, n, 5, steps=25, plane=True, kernel=gpu.KERNEL_TILED, expect="kernel=tiled")


@pytest.mark.parametrize("model_name,kw", [("DNKG", dict(m=1.0, beta=0.3)), ("FPUT", dict(alpha=0.3, beta=0.0)),
                                           ("FPUT", dict(alpha=0.25, beta=1.0))])
def test_move_tiled_force_forms(gpu, model_name, kw):
    check_move(gpu, model_name, 17000, 2, steps=20, kw=kw, dens=0.1, expect="kernel=tiled")


def test_autotune_long_chain_includes_tiled(gpu):
    N = gpu
    model, kw, masses, s0 = make_case(N, "DNKG", 20000, 4, 3)
    with N.Ensemble(model, 20000, 4, 0.05, **kw) as ens:
        ens.upload(s0)
        report = ens.autotune(20)
        desc = ens.describe()
        banner(f"autotune n=20000: {report} -> {desc}")
        assert "tiled" in report and "split" in report
        ens.move(30)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0, 30, 0.05)
    assert chain_rel_err(got, want) <= STATE_TOL


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n", [3, 5, 7, 12, 17, 31, 33, 45, 63])
def test_move_short_chains(gpu, model_name, n):
    """Short chains whose length is not a power of two: a thread per chain up
    to 32 sites, 4 lanes with one site fewer on the last lanes above."""
    banner(f"short chains {model_name} n={n}")
    model, kw, masses, s0 = make_case(gpu, model_name, n, 70, 6)
    with gpu.Ensemble(model, n, 70, 0.05, masses=masses, **kw) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(100)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0, 100, 0.05)
    err = chain_rel_err(got, want)
    banner(f"  {desc}: rel err {err:.3e}")
    assert "kernel=wave" in desc and f"lanes_per_chain={1 if n <= 32 else 4}" in desc
    assert np.isfinite(got).all() and err <= STATE_TOL, (desc, err)


GROUP_CASES = [(m, n, c) for n, c in [(2500, 5), (3001, 3), (4096, 4), (6000, 3), (8192, 3), (16384, 2)]
               for m in ("DNKG", "FPUT", "Toda", "dDNKG") if m != "dDNKG" or n <= 4096]


@pytest.mark.parametrize("model_name,n,copies", GROUP_CASES)
def test_move_group_default(gpu, model_name, n, copies):
    """Chains longer than a wavefront holds (n > 2048) run on the
    workgroup-resident kernel by default, with 2, 4 or 8 wavefronts per chain."""
    banner(f"group move {model_name} n={n} copies={copies}")
    model, kw, masses, s0 = make_case(gpu, model_name, n, copies, 4)
    with gpu.Ensemble(model, n, copies, 0.05, masses=masses, **kw) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(60)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0, 60, 0.05)
    err = chain_rel_err(got, want)
    banner(f"  {desc}: rel err {err:.3e}")
    assert "kernel=group" in desc
    assert any(f"lanes_per_chain={L} " in desc for L in (128, 256, 512)), desc
    assert np.isfinite(got).all() and err <= STATE_TOL, (desc, err)


@pytest.mark.parametrize("n,L,S", [(4096, 128, 32), (4096, 512, 8), (3000, 192, 16), (2048, 128, 16),
                                   (1024, 128, 8), (16384, 512, 32), (5000, 320, 16), (2100, 192, 11)])
def test_move_group_shapes(gpu, n, L, S):
    for model_name in ("DNKG", "Toda"):
        check_move(gpu, model_name, n, 3, steps=40, lanes_per_chain=L, sites_per_lane=S)


def test_move_group_rejects_invalid_shapes(gpu):
    for n, L, S in [(4096, 192, 16), (4096, 1024, 4), (4096, 256, 8), (100000, 512, 32)]:
        with pytest.raises(gpu.NlcError):
            gpu.Ensemble(gpu.DNKG, n, 2, 0.05, m=1.0, beta=1.0, lanes_per_chain=L, sites_per_lane=S)


def test_autotune_includes_group_shapes(gpu):
    N = gpu
    model, kw, masses, s0 = make_case(N, "FPUT", 4096, 64, 3)
    with N.Ensemble(model, 4096, 64, 0.05, **kw) as ens:
        ens.upload(s0)
        report = ens.autotune(5)
        banner(f"autotune n=4096: {report} -> {ens.describe()}")
        assert "256x16" in report and "128x32" in report and "split" in report
        assert np.array_equal(ens.download(), s0)
        ens.move(30)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0[:4], 30, 0.05)
    assert chain_rel_err(got[:4], want) <= STATE_TOL


@pytest.mark.parametrize("n", [36, 40, 60, 192, 960, 1984])
def test_exact_multiples_stay_on_the_wave_kernel(gpu, n):
    """Lengths n = 4 S or 64 S with S not a power of two have no exact-fit
    instance; they run on the ragged-lane one with every lane full, by default
    and among the autotuner's candidates."""
    N = gpu
    model, kw, masses, s0 = make_case(N, "DNKG", n, 24, 5)
    with N.Ensemble(model, n, 24, 0.05, **kw) as ens:
        assert "kernel=wave" in ens.describe(), ens.describe()
        ens.upload(s0)
        report = ens.autotune(5)
        desc = ens.describe()
        banner(f"n={n}: {report} -> {desc}")
        L = 4 if n < 64 else 64
        assert f" {L}x{n // L} " in report, report  # a candidate of the autotuner
        assert "kernel=split" not in desc, (report, desc)
        ens.move(40)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0, 40, 0.05)
    assert chain_rel_err(got, want) <= STATE_TOL


@pytest.mark.parametrize("n,copies,shape", [(64, 65536, (4, 16)), (128, 32768, (8, 16)), (256, 16384, (16, 16)),
                                            (64, 96, (64, 1)), (64, 4096, (64, 1)), (128, 2048, (64, 2)),
                                            (256, 8192, (32, 8)), (32, 65536, (4, 8)), (32, 8192, (32, 1)),
                                            (16, 96, (16, 1))])
def test_default_shape_short_powers_of_two(gpu, n, copies, shape):
    """Powers of two 64..256 default to 16 sites per lane, traded for lanes
    while the ensemble has fewer than 4096 wavefronts."""
    N = gpu
    model, kw, masses, s0 = make_case(N, "FPUT", n, copies, 6)
    with N.Ensemble(model, n, copies, 0.05, **kw) as ens:
        desc = ens.describe()
        assert f"lanes_per_chain={shape[0]} sites_per_lane={shape[1]} " in desc, desc
        ens.upload(s0)
        ens.move(20)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0[:16], 20, 0.05)
    assert chain_rel_err(got[:16], want) <= STATE_TOL


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
@pytest.mark.parametrize("n", [65, 80, 100, 333, 511])
def test_move_16_lane_ragged(gpu, model_name, n):
    """16 lanes per chain (four chains to a wavefront, 9 copies: the last
    wavefront partly filled) with S = ceil(n/16) sites per lane."""
    check_move(gpu, model_name, n, 9, steps=40, lanes_per_chain=16, sites_per_lane=-(-n // 16))


@pytest.mark.parametrize("n,copies,shape", [(300, 16384, (16, 19)), (100, 20000, (16, 7)), (300, 100, (64, 5)),
                                            (500, 16384, (64, 8))])
def test_default_shape_16_lanes(gpu, n, copies, shape):
    """Lengths 65..384 that are not powers of two default to 16 lanes when the
    ensemble fills 4096 wavefronts, else (and above 384 sites) to 64 lanes."""
    N = gpu
    model, kw, masses, s0 = make_case(N, "DNKG", n, copies, 7)
    with N.Ensemble(model, n, copies, 0.05, **kw) as ens:
        desc = ens.describe()
        assert f"lanes_per_chain={shape[0]} sites_per_lane={shape[1]} exact=0" in desc, desc
        ens.upload(s0)
        ens.move(20)
        got = ens.download()
    want = oracle_move(model, kw, masses, s0[-8:], 20, 0.05)
    assert chain_rel_err(got[-8:], want) <= STATE_TOL


def test_move_single_copy_and_single_step(gpu):
    check_move(gpu, "DNKG", 1024, 1, steps=1)
    check_move(gpu, "FPUT", 64, 1, steps=1)


def test_move_batches_compose(gpu):
    """b + b steps in one call equals two calls of b (the merged c7/c0 drift
    is undone at the end of every batch), as the reference's resume semantics need."""
    N = gpu
    model, kw, masses, s0 = make_case(N, "DNKG", 1024, 8, 3)
    with N.Ensemble(model, 1024, 8, 0.05, **kw) as a, N.Ensemble(model, 1024, 8, 0.05, **kw) as b:
        a.upload(s0)
        b.upload(s0)
        a.move(40)
        b.move(20)
        b.move(20)
        assert chain_rel_err(a.download(), b.download()) < 1e-13


@pytest.mark.parametrize("shape", [dict(), dict(lanes_per_chain=2, sites_per_lane=32)])
def test_toda_far_stretched_bond_gives_nonfinite_state(gpu, shape):
    """A Toda bond stretched by 1e9 or 1e12 (exponent arguments of 9e10 and
    9e13 in the kernel's units, beyond the 2^30 where a wrapped integer part
    made the exponential finite) gives an infinite force and a non-finite chain
    after one step, in both exponential forms (the table form on the default
    64 x 1 shape, the table-free one on 2 x 32); the other chains of the
    ensemble stay on the oracle's trajectory."""
    N = gpu
    n, copies = 64, 6
    model, kw, masses, s0 = make_case(N, "Toda", n, copies, 8)
    s0 = s0.copy()
    s0[1, 20, 0] += 1e9
    s0[3, 41, 0] -= 1e12
    with N.Ensemble(model, n, copies, 0.05, **kw, **shape) as ens:
        ens.upload(s0)
        ens.move(1)
        got = ens.download()
        desc = ens.describe()
    banner(f"  {desc}: finite chains {[bool(np.isfinite(got[c]).all()) for c in range(copies)]}")
    good = [0, 2, 4, 5]
    assert not np.isfinite(got[1]).all() and not np.isfinite(got[3]).all()
    assert np.isfinite(got[good]).all()
    want = oracle_move(model, kw, masses, s0[good], 1, 0.05)
    assert chain_rel_err(got[good], want) <= STATE_TOL


def test_move_is_deterministic(gpu):
    N = gpu
    model, kw, masses, s0 = make_case(N, "Toda", 512, 64, 4)
    outs = []
    for _ in range(2):
        with N.Ensemble(model, 512, 64, 0.05, **kw) as e:
            e.upload(s0)
            e.move(30)
            outs.append(e.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n", [1024, 4096])
def test_eigensystem_device_path(gpu, n):
    """Above 512 sites the eigensystem runs on the device (rocSOLVER dsyevd):
    eigenvalues against LAPACK, the eigen-equation and orthonormality."""
    masses = gpu.ic_masses(n, 0.5, 1.5, 7)
    t0 = time.perf_counter()
    om, V = gpu.eigensystem(masses + 2)
    banner(f"eigensystem n={n}: {time.perf_counter() - t0:.2f} s")
    A = np.diag(masses + 2) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    lam = np.linalg.eigvalsh(A)
    np.testing.assert_allclose(om, np.sqrt(lam), rtol=1e-12)
    assert np.abs(A @ V - V * om ** 2).max() <= 1e-11 * np.abs(lam).max()
    assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-12


def test_ensemble_beyond_2_30_sites(gpu):
    """FPUT 64 sites x 4096*4097 copies = 1.07e9 sites (17 GB of state, past
    2^30), built as 4097 tiles of one seeded 4096-copy ensemble: every tile
    ends bit-identical, sampled chains match the oracle, the chain energies of
    all tiles agree, and the mode energies are the tile's times 4097."""
    N = gpu
    n, base_copies, tiles, steps = 64, 4096, 4097, 20
    model, kw, masses, base = make_case(N, "FPUT", n, base_copies, 8)
    copies = base_copies * tiles
    assert n * copies > 1 << 30
    banner(f"large ensemble: {n} x {copies} = {n * copies:.3e} sites")
    s0 = np.tile(base, (tiles, 1, 1))
    with N.Ensemble(model, n, copies, 0.1, **kw) as ens:
        ens.upload(s0)
        del s0
        ens.move(steps)
        got = ens.download().reshape(tiles, base_copies, n, 2)
        E, _ = ens.chain_energies()
        raw = ens.linenergies_raw()
    first = got[0]
    for t in range(1, tiles):
        assert np.array_equal(got[t], first), t
    del got
    want = oracle_move(model, kw, masses, base[:8], steps, 0.1)
    assert chain_rel_err(first[:8], want) <= STATE_TOL
    E = E.reshape(tiles, base_copies)
    assert np.array_equal(E, np.broadcast_to(E[0], E.shape))
    np.testing.assert_allclose(E[0, :8], O.chain_energy(O.FPUT, want, **kw), rtol=1e-12)
    e_tile = O.linenergies_real(first, O.dispersion_real(n, 0.0))
    assert vec_rel_err(raw, tiles * e_tile) <= 1e-11
    with pytest.raises(N.NlcError):
        N.Ensemble(N.DNLS, 2, 1 << 24, 0.01, beta=1.0)  # one workgroup per DNLS chain: < 2^24 copies


# ------------------------------------------------------------ energy pass

@pytest.mark.parametrize("model_name,n,copies", [("DNKG", 64, 50), ("DNKG", 1024, 64), ("FPUT", 256, 33),
                                                 ("Toda", 128, 20), ("dDNKG", 512, 70), ("dDNKG", 100, 9),
                                                 ("DNKG", 3000, 3), ("FPUT", 96, 5), ("DNKG", 7, 40),
                                                 # 8192-point LDS FFT; mixed-radix FFT (5120 = 2^10 5,
                                                 # 6000 = 2^4 3 5^3, 7000 = 2^3 5^3 7); DFT in LDS (3001)
                                                 ("DNKG", 8192, 5), ("Toda", 5120, 3), ("FPUT", 6000, 3),
                                                 ("DNKG", 7000, 2), ("Toda", 3001, 3),
                                                 # short chains, 1024/n per workgroup: partial groups,
                                                 # several rounds per workgroup
                                                 ("DNKG", 8, 129), ("FPUT", 16, 3), ("Toda", 32, 1000),
                                                 ("FPUT", 64, 70000),
                                                 # two-level FFT above 8192 sites (600 copies of 65536
                                                 # sites: two chunks of the 2^25-site scratch)
                                                 ("DNKG", 16384, 3), ("Toda", 32768, 2), ("FPUT", 131072, 2),
                                                 ("DNKG", 65536, 600),
                                                 # Bluestein through the two-level FFT (L = 16384, 65536,
                                                 # 131072; 2100 copies of 6007 sites: two chunks)
                                                 ("DNKG", 6007, 3), ("FPUT", 20001, 2), ("Toda", 33001, 1)])
def test_linenergies_and_entropy(gpu, model_name, n, copies):
    N = gpu
    banner(f"linenergies {model_name} n={n} copies={copies}")
    if n > 8192:  # plane-wave starts cost O(sites x modes) to generate: a seeded random state instead
        model, kw, masses = N.MODELS[model_name], dict(MODEL_KW[model_name]), None
        s0 = 0.1 * np.random.default_rng(9).normal(size=(copies, n, 2))
    else:
        model, kw, masses, s0 = make_case(N, model_name, n, copies, 9)
        s0 = s0 + 0.01 * np.random.default_rng(1).normal(size=s0.shape)  # populate every mode
    with N.Ensemble(model, n, copies, 0.05, masses=masses, **kw) as ens:
        ens.upload(s0)
        e_gpu, s_gpu = ens.linenergies()
        raw_gpu = ens.linenergies_raw()
    if model == N.DDNKG:
        om, V = N.eigensystem(masses + 2)
        raw = O.linenergies_ddnkg(s0, V, om)
        norm, a0is0 = 0.5 / copies, False
    else:
        raw = O.linenergies_real(s0, O.dispersion_real(n, kw.get("m", 0.0)))
        norm, a0is0 = 0.5 / copies / n, kw.get("m", 0.0) == 0.0
    e = O.calc_linenergies(raw, norm, a0is0)
    wt, inf = O.entropies(e, a0is0)
    banner(f"  e rel {vec_rel_err(e_gpu, e):.3e}  S_WT {s_gpu[0]:.12g}/{wt:.12g}  S_INF {s_gpu[1]:.12g}/{inf:.12g}")
    assert vec_rel_err(raw_gpu, raw) <= ENERGY_TOL
    assert vec_rel_err(e_gpu, e) <= ENERGY_TOL
    assert abs(s_gpu[0] - wt) <= ENERGY_TOL * max(1, abs(wt))
    assert abs(s_gpu[1] - inf) <= ENERGY_TOL * max(1, abs(inf))


@pytest.mark.parametrize("n,tiles", [(6007, 700), (65536, 200)])
def test_linenergies_long_chains_in_chunks(gpu, n, tiles):
    """More copies than one scratch chunk holds (2^25 values: 2048 chains of
    6007 sites through Bluestein's L = 16384, 512 chains of 65536): three
    seeded chains tiled, so the raw mode energies are the tile count times
    the oracle's for the three."""
    N = gpu
    base = 0.1 * np.random.default_rng(11).normal(size=(3, n, 2))
    with N.Ensemble(N.DNKG, n, 3 * tiles, 0.05, m=1.0, beta=1.0) as ens:
        ens.upload(np.tile(base, (tiles, 1, 1)))
        raw = ens.linenergies_raw()
    want = tiles * O.linenergies_real(base, O.dispersion_real(n, 1.0))
    assert vec_rel_err(raw, want) <= 1e-11


def test_entropy_mask_on_gpu(gpu):
    N = gpu
    n, copies = 128, 16
    model, kw, masses, s0 = make_case(N, "FPUT", n, copies, 5)
    s0 = s0 + 0.01 * np.random.default_rng(2).normal(size=s0.shape)
    mask = np.zeros(n, np.uint8)
    mask[[0, 1, 2, 5, 9, 40, 77]] = 1
    with N.Ensemble(model, n, copies, 0.05, **kw) as ens:
        ens.upload(s0)
        ens.set_entropy_mask(mask, True)
        e_gpu, s_gpu = ens.linenergies()
        with pytest.raises(N.NlcError):
            ens.set_entropy_mask(np.eye(1, n, 0, dtype=np.uint8)[0] + np.eye(1, n, 1, dtype=np.uint8)[0], True)
    e = O.calc_linenergies(O.linenergies_real(s0, O.dispersion_real(n, 0.0)), 0.5 / copies / n, True)
    idx = O.entropy_modes(n, mask, True)
    wt, inf = O.entropies(e, True, idx)
    assert abs(s_gpu[0] - wt) <= ENERGY_TOL * abs(wt) and abs(s_gpu[1] - inf) <= ENERGY_TOL * abs(inf)


@pytest.mark.parametrize("model_name", ["DNKG", "FPUT", "Toda", "dDNKG"])
def test_chain_energies(gpu, model_name):
    N = gpu
    model, kw, masses, s0 = make_case(N, model_name, 256, 12, 6)
    with N.Ensemble(model, 256, 12, 0.05, masses=masses, **kw) as ens:
        ens.upload(s0)
        E, _ = ens.chain_energies()
    want = O.chain_energy(model, s0, masses=masses, **kw)
    np.testing.assert_allclose(E, want, rtol=1e-12)


# ------------------------------------------------ full-size property checks

def test_full_size_dnkg_sampled_parity_and_drift(gpu):
    """At the benchmark shape (1024 sites x 16384 copies): chains are
    independent, so a sample of copies re-run on the oracle checks parity at
    full size; energy drift over 1000 steps and run-to-run determinism are
    size-independent properties."""
    N = gpu
    n, copies, dt, steps = 1024, 16384, 0.05, 200
    banner("full-size DNKG 1024 x 16384")
    s0 = N.ic_plane_waves(n, copies, 1.0, 0, n // 8, 0.5, 2)
    with N.Ensemble(N.DNKG, n, copies, dt, m=1.0, beta=1.0) as ens:
        ens.upload(s0)
        E0, _ = ens.chain_energies()
        ens.move(steps)
        got = ens.download()
        E1, _ = ens.chain_energies()
        e_lin, ent = ens.linenergies()
    pick = np.sort(np.random.default_rng(0).choice(copies, 256, replace=False))
    want = oracle_parallel(O.move_real, s0[pick], model=O.DNKG, steps=steps, dt=dt, m=1.0, beta=1.0)
    err = chain_errors(got[pick], want)
    banner(f"  256 sampled chains: largest per-chain rel err {err.max():.3e}, median {np.median(err):.3e}")
    assert err.max() <= STATE_TOL
    drift = np.mean(np.abs(E1 - E0) / E0)
    banner(f"  sampled parity ok, mean relative energy error {drift:.3e}, entropies {ent}")
    # The energy error at this dt is the integrator's truncation error (same
    # method on both sides): the GPU's must equal the oracle's chain by chain.
    E0_o = O.chain_energy(O.DNKG, s0[pick], m=1.0, beta=1.0)
    E1_o = O.chain_energy(O.DNKG, want, m=1.0, beta=1.0)
    np.testing.assert_allclose((E1[pick] - E0[pick]) / E0[pick], (E1_o - E0_o) / E0_o, rtol=0, atol=1e-12)
    assert drift < 1e-6
    assert np.isfinite(e_lin).all() and np.isfinite(ent).all()


# ------------------------------------------------ dDNKG projection forms

@pytest.mark.parametrize("proj", ["mfma", "mfma_wide", "valu"])
@pytest.mark.parametrize("n,copies", [(512, 70), (100, 9), (130, 129), (48, 200), (16, 64)])
def test_ddnkg_projection_forms(gpu, proj, n, copies):
    """Both forms of the fused projection (FP64 matrix cores, FP64 vector ALU)
    against the oracle's O(copies n^2) projection, including chain lengths that
    are not multiples of the tile sizes and partial copy tiles."""
    N = gpu
    banner(f"dDNKG projection {proj} n={n} copies={copies}")
    model, kw, masses, s0 = make_case(N, "dDNKG", n, copies, 11)
    s0 = s0 + 0.01 * np.random.default_rng(3).normal(size=s0.shape)
    p = {"mfma": N.PROJ_MFMA_F64, "mfma_wide": N.PROJ_MFMA_F64_WIDE, "valu": N.PROJ_VALU_F64}[proj]
    with N.Ensemble(model, n, copies, 0.05, masses=masses, projection=p, **kw) as ens:
        ens.upload(s0)
        raw_gpu = ens.linenergies_raw()
        assert f"projection={proj}_f64" in ens.describe()
    om, V = N.eigensystem(masses + 2)
    raw = O.linenergies_ddnkg(s0, V, om)
    assert vec_rel_err(raw_gpu, raw) <= ENERGY_TOL


def test_ddnkg_projection_full_size(gpu):
    """The dDNKG workload's shape (512 sites x 8192 copies): the two forms agree,
    and the modal sum equals the harmonic energy sum_c (p.p + q.A.q) with
    A = diag(m_x + 2) - periodic adjacency, since V is orthonormal."""
    N = gpu
    n, copies = 512, 8192
    masses = N.ic_masses(n, 0.5, 1.5, 5)
    om, V = N.eigensystem(masses + 2)
    s0 = N.ic_eigenmodes(n, copies, om, V, n // 8, 0.5, 5)
    s0 = s0 + 0.01 * np.random.default_rng(4).normal(size=s0.shape)
    raws = {}
    for name, p in (("mfma", N.PROJ_MFMA_F64), ("mfma_wide", N.PROJ_MFMA_F64_WIDE), ("valu", N.PROJ_VALU_F64)):
        with N.Ensemble(N.DDNKG, n, copies, 0.05, masses=masses, beta=1.0, projection=p) as ens:
            ens.upload(s0)
            raws[name] = ens.linenergies_raw()
    q, pm = s0[..., 0], s0[..., 1]
    harmonic = (pm * pm).sum() + ((masses + 2) * q * q).sum() - 2.0 * (q * np.roll(q, -1, axis=1)).sum()
    banner(f"  mfma vs valu {vec_rel_err(raws['mfma'], raws['valu']):.3e}; "
           f"sum {raws['mfma'].sum():.15g} vs harmonic {harmonic:.15g}")
    assert vec_rel_err(raws["mfma"], raws["valu"]) <= 1e-12
    assert vec_rel_err(raws["mfma_wide"], raws["valu"]) <= 1e-12
    assert abs(raws["mfma"].sum() - harmonic) <= 1e-11 * harmonic


@pytest.mark.parametrize("n,copies", [(32, 5), (512, 70), (96, 130), (64, 8192), (320, 200), (512, 8192)])
def test_ddnkg_projection_split_bf16(gpu, n, copies):
    """BF16 matrix cores with three-term operand splits: mode energies within
    1e-6 of the largest mode energy of the FP64 result (measured: about 1e-7),
    modes above 1e-3 of the largest within 1e-5 relative, entropies within 1e-6."""
    N = gpu
    model, kw, masses, s0 = make_case(N, "dDNKG", n, copies, 13)
    s0 = s0 + 0.01 * np.random.default_rng(6).normal(size=s0.shape)
    with N.Ensemble(model, n, copies, 0.05, masses=masses, projection=N.PROJ_SPLIT_BF16, **kw) as ens:
        assert "projection=split_bf16" in ens.describe()
        ens.upload(s0)
        raw = ens.linenergies_raw()
        e_gpu, s_gpu = ens.linenergies()
    om, V = N.eigensystem(masses + 2)
    ref = O.linenergies_ddnkg(s0, V, om)
    abs_err = np.abs(raw - ref).max() / ref.max()
    big = ref > 1e-3 * ref.max()
    rel_big = (np.abs(raw - ref)[big] / ref[big]).max()
    e = O.calc_linenergies(ref, 0.5 / copies, False)
    wt, inf = O.entropies(e, False)
    banner(f"  split BF16 n={n} copies={copies}: max err / max energy {abs_err:.2e}, "
           f"rel err of modes > 1e-3 max {rel_big:.2e}, S_WT {abs(s_gpu[0] - wt) / abs(wt):.2e}")
    assert abs_err <= 1e-6 and rel_big <= 1e-5
    assert abs(s_gpu[0] - wt) <= 1e-6 * abs(wt) and abs(s_gpu[1] - inf) <= 1e-6 * abs(inf)


def test_ddnkg_split_bf16_needs_multiple_of_32(gpu):
    N = gpu
    with pytest.raises(N.NlcError):
        N.Ensemble(N.DDNKG, 100, 4, 0.05, masses=np.ones(100), beta=1.0, projection=N.PROJ_SPLIT_BF16)


def test_ddnkg_autotune_reports_projection(gpu):
    N = gpu
    model, kw, masses, s0 = make_case(N, "dDNKG", 256, 300, 12)
    with N.Ensemble(model, 256, 300, 0.05, masses=masses, **kw) as ens:
        ens.upload(s0)
        before = ens.linenergies_raw()
        report = ens.autotune(5)
        banner(f"autotune: {report}")
        assert "projection: mfma" in report and "valu" in report
        after = ens.linenergies_raw()
        assert np.array_equal(ens.download(), s0)
    assert vec_rel_err(after, before) <= 1e-13
