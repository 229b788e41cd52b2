"""The FP32-compensated DNLS step (NLC_FP32C, 4096-site chains) against the
FP64 oracle and the FP64 kernel, on the DNLS workload's parameters (beta = 1,
dt = 0.01, norm density 1 on |k| < 256).

The FP32C step is a different rounding of the same method, so it is checked
against bounds on its distance from the FP64 trajectory (chain by chain) and
on the drift of the invariants the FP64 method conserves.  Measured on MI355X
(seed 3): after 100 steps the largest per-chain error against the FP64 oracle
is 6.4e-7 (median 4.8e-7), the norm drift 8e-10, linear energies 1.1e-7 of the
largest mode energy, entropies 5e-8; over 1000 steps of 4096 chains the norm
drifts by 1.9e-8 on average (3.6e-8 at most; FP64: 1.6e-12) and the energy by
6.4e-8 (1.3e-7 at most; FP64: 1.7e-12).  The bounds below are those values with
a margin of about 3 to 5:
  * psi after 100 steps within STATE_TOL of the FP64 oracle, per chain;
  * norm of every chain conserved to NORM_TOL_100 over 100 steps and to
    NORM_TOL over 1000;
  * linear energies within ENERGY_TOL of the largest mode energy of the FP64
    oracle's, entropies within ENERGY_TOL relative, after 100 steps;
  * the relative energy error over 1000 steps within DRIFT_TOL (mean) and
    3 DRIFT_TOL (largest).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors, vec_rel_err

pytestmark = pytest.mark.gpu

N_SITES, DT, BETA = 4096, 0.01, 1.0
STATE_TOL = 3e-6
NORM_TOL_100 = 5e-9
NORM_TOL = 1.5e-7
ENERGY_TOL = 5e-7
DRIFT_TOL = 3e-7


def banner(msg):
    print(f"[gpu-dnls-fp32c] {msg}", flush=True)
    sys.stdout.flush()


def ensemble(N, copies, precision):
    return N.Ensemble(N.DNLS, N_SITES, copies, DT, beta=BETA, precision=precision)


def test_fp32c_parity_against_fp64_oracle(gpu):
    N = gpu
    copies, steps = 8, 100
    s0 = N.ic_dnls(N_SITES, copies, 256, 1.0, 3)
    with ensemble(N, copies, N.FP32C) as ens:
        desc = ens.describe()
        ens.upload(s0)
        _, n0 = ens.chain_energies()
        ens.move(steps)
        got = ens.download()
        E1, n1 = ens.chain_energies()
        e_gpu, s_gpu = ens.linenergies()
    banner(desc)
    assert "fp32c" in desc
    want = O.move_dnls(s0, steps, DT, BETA)
    err = chain_errors(got, want)
    norm = np.abs(n1 / n0 - 1)
    e = O.calc_linenergies(O.linenergies_dnls(want), 1.0 / (N_SITES * copies), True)
    wt, inf = O.entropies(e, True)
    e_err = vec_rel_err(e_gpu, e)
    s_err = max(abs(s_gpu[0] - wt) / abs(wt), abs(s_gpu[1] - inf) / abs(inf))
    E_err = np.abs(E1 / O.chain_energy(O.DNLS, want, beta=BETA) - 1).max()
    banner(f"  100 steps: per-chain rel err max {err.max():.3e} median {np.median(err):.3e}; norm drift "
           f"{norm.max():.3e}; linenergies {e_err:.3e}; entropies {s_err:.3e}; chain energy vs oracle {E_err:.3e}")
    assert np.isfinite(got).all()
    assert err.max() <= STATE_TOL
    assert norm.max() <= NORM_TOL_100
    assert e_err <= ENERGY_TOL and s_err <= ENERGY_TOL
    assert E_err <= ENERGY_TOL


def test_fp32c_full_size_drift(gpu):
    """4096 x 4096 (the DNLS workload), one 1000-step batch in each precision
    from the same state: norm drift of every chain, and the mean relative
    energy error next to the FP64 kernel's."""
    N = gpu
    copies, steps = 4096, 1000
    s0 = N.ic_dnls(N_SITES, copies, 256, 1.0, 3)
    res = {}
    for prec in (N.FP64, N.FP32C):
        with ensemble(N, copies, prec) as ens:
            ens.upload(s0)
            E0, n0 = ens.chain_energies()
            ens.move(steps)
            E1, n1 = ens.chain_energies()
            fin = bool(np.isfinite(ens.download()).all())
        res[prec] = (np.abs(n1 / n0 - 1), np.abs(E1 / E0 - 1), fin)
    (n64, e64, f64), (n32, e32, f32) = res[N.FP64], res[N.FP32C]
    banner(f"  1000 steps: norm drift FP64 max {n64.max():.2e}, FP32C max {n32.max():.2e} mean {n32.mean():.2e}; "
           f"energy error FP64 mean {e64.mean():.2e}, FP32C mean {e32.mean():.2e} max {e32.max():.2e}")
    assert f64 and f32
    assert n64.max() <= 1e-11 and e64.mean() <= 1e-10
    assert n32.max() <= NORM_TOL
    assert e32.mean() <= DRIFT_TOL and e32.max() <= 3 * DRIFT_TOL


def test_fp32c_sampled_chains_full_size(gpu):
    """The full ensemble advanced 50 steps in FP32C; 32 chains re-run on the
    FP64 oracle."""
    N = gpu
    copies, steps = 4096, 50
    s0 = N.ic_dnls(N_SITES, copies, 256, 1.0, 9)
    with ensemble(N, copies, N.FP32C) as ens:
        ens.upload(s0)
        ens.move(steps)
        got = ens.download()
    pick = np.sort(np.random.default_rng(4).choice(copies, 32, replace=False))
    from parity import oracle_parallel
    want = oracle_parallel(O.move_dnls, s0[pick], steps=steps, dt=DT, beta=BETA)
    err = chain_errors(got[pick], want)
    banner(f"  50 steps, 32 sampled chains: per-chain rel err max {err.max():.3e}")
    assert err.max() <= STATE_TOL


def test_fp32c_batches_and_graph(gpu):
    """Two batches of 10 steps against one of 20 (the split closes a batch with
    the FP64 c_7 rotation and reloads the state into hi + lo: a rounding-level
    difference), and graph replays (nlc_batch) identical to the eager path."""
    N = gpu
    s0 = N.ic_dnls(N_SITES, 4, 256, 1.0, 5)
    with ensemble(N, 4, N.FP32C) as a, ensemble(N, 4, N.FP32C) as b:
        a.upload(s0)
        a.move(10)
        a.move(10)
        b.upload(s0)
        b.move(20)
        d = chain_errors(a.download(), b.download())
        banner(f"  10 + 10 vs 20 steps: {d.max():.3e}")
        assert d.max() <= 1e-7
        a.upload(s0)
        b.upload(s0)
        for rep in range(3):
            eg, sg = a.batch(10)
            b.move(10)
            ee, se = b.linenergies()
            assert np.array_equal(a.download(), b.download()), rep
            assert np.array_equal(eg, ee) and np.array_equal(sg, se), rep


def test_fp32c_unsupported_combinations(gpu):
    N = gpu
    # (FPUT and Toda FP32C: power-of-two chains up to 2048 sites, tests/test_gpu_fp32c_real.py)
    for model, n, kw in ((N.DNKG, 1024, dict(m=1.0, beta=1.0)), (N.TODA, 4096, dict(alpha=1.0)),
                         (N.DNLS, 1024, dict(beta=1.0)), (N.DNLS, 8192, dict(beta=1.0))):
        with pytest.raises(N.NlcError) as ex:
            N.Ensemble(model, n, 4, 0.01, precision=N.FP32C, **kw)
        assert ex.value.code == N.NLC_EUNSUPPORTED
    with pytest.raises(N.NlcError):
        N.Ensemble(N.DNLS, N_SITES, 4, 0.01, beta=1.0, precision=N.FP32C, sites_per_lane=8)
    with pytest.raises(N.NlcError) as ex:
        N.Ensemble(N.DNLS, N_SITES, 4, 0.01, beta=1.0, precision=7)
    assert ex.value.code == N.NLC_EARG
