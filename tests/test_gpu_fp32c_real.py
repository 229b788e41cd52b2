"""The FP32-compensated step of FPUT and Toda (NLC_FP32C: q and p on chip as
FP32 pairs with Kahan-compensated updates; FPUT forces in FP32, the Toda
exponential in FP64; state in HBM, energies and entropies in FP64) against the
FP64 oracle and the FP64 kernels.

FP32C is a different rounding of the same method, so it is checked against
bounds on its distance from the FP64 trajectory, chain by chain, and on the
energy drift next to the FP64 kernel's.  Measured on MI355X (100 steps,
dt = 0.05, seeded plane waves): the largest per-chain state error against the
FP64 oracle is 1.4e-6 (2-site chains; 1e-7 to 7e-7 from 32 sites on), chain
energies within 4.5e-7 of the oracle's; over 1000 steps of 1024-site chains
the mean energy drift equals the FP64 kernel's to 2 % (FPUT 1.55e-7 against
1.53e-7, Toda 3.15e-7 against 3.18e-7: the integrator's own error at this dt
dominates); ragged and mixed ensembles stay within 2e-7 of the FP64 kernels
after 100 steps.  The bounds below are those values with a margin of 3 to 5.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors

pytestmark = pytest.mark.gpu

STATE_TOL = 5e-6      # per chain, 100 steps, against the FP64 oracle
ENERGY_TOL = 2e-6     # chain energies after 100 steps, against the oracle's
RAGGED_TOL = 1e-6     # per chain, 100 steps, against the FP64 kernel / oracle
ORACLE_MODEL = {1: O.FPUT, 2: O.TODA}

CASES = {
    "fput_cubic": (1, dict(beta=1.0), 0.5),
    "fput_plain": (1, dict(alpha=0.25, beta=1.0), 0.5),
    # alpha-FPUT alone has a potential unbounded below: a low energy density keeps it bound
    "fput_quad": (1, dict(alpha=0.3), 0.05),
    "toda": (2, dict(alpha=1.0), 0.25),
}


def banner(msg):
    print(f"[gpu-fp32c-real] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("n,copies,shape", [
    (64, 256, {}), (1024, 32, {}), (32, 64, {}), (2048, 8, {}), (2, 64, {}),
    (64, 40, dict(lanes_per_chain=8, sites_per_lane=8)), (256, 16, dict(lanes_per_chain=32, sites_per_lane=8)),
    (64, 16, dict(lanes_per_chain=2, sites_per_lane=32)), (16, 70, dict(lanes_per_chain=1, sites_per_lane=16)),
])
def test_fp32c_parity_against_fp64_oracle(gpu, case, n, copies, shape):
    N = gpu
    model, kw, dens = CASES[case]
    s0 = N.ic_plane_waves(n, copies, 0.0, 1, max(2, n // 8), dens, 3)
    with N.Ensemble(model, n, copies, 0.05, precision=N.FP32C, **kw, **shape) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(100)
        got = ens.download()
        E, _ = ens.chain_energies()
    want = O.move_real(ORACLE_MODEL[model], s0, 100, 0.05, threads=8, **kw)
    err = chain_errors(got, want)
    e_err = np.abs(E / O.chain_energy(ORACLE_MODEL[model], want, **kw) - 1).max()
    banner(f"{desc}: per-chain err max {err.max():.2e} median {np.median(err):.2e}; chain energies {e_err:.2e}")
    assert "precision=fp32c" in desc and "exact=1" in desc
    if shape:
        assert f"lanes_per_chain={shape['lanes_per_chain']} sites_per_lane={shape['sites_per_lane']}" in desc
    assert np.isfinite(got).all()
    assert err.max() <= STATE_TOL
    assert e_err <= ENERGY_TOL


@pytest.mark.parametrize("case", ["fput_cubic", "toda"])
def test_fp32c_drift_next_to_fp64(gpu, case):
    """1000 steps of 1024 x 256 from one state in both precisions: the FP32C
    energy drift is the FP64 kernel's (the method's truncation error at
    dt = 0.05 is far above the FP32C rounding)."""
    N = gpu
    model, kw, dens = CASES[case]
    s0 = N.ic_plane_waves(1024, 256, 0.0, 1, 128, dens, 5)
    drift = {}
    for prec in (N.FP64, N.FP32C):
        with N.Ensemble(model, 1024, 256, 0.05, precision=prec, **kw) as ens:
            ens.upload(s0)
            E0, _ = ens.chain_energies()
            ens.move(1000)
            E1, _ = ens.chain_energies()
            assert np.isfinite(ens.download()).all()
        drift[prec] = np.abs(E1 / E0 - 1)
    d64, d32 = drift[N.FP64], drift[N.FP32C]
    banner(f"{case} 1000 steps: energy drift FP64 mean {d64.mean():.2e} max {d64.max():.2e}; "
           f"FP32C mean {d32.mean():.2e} max {d32.max():.2e}")
    assert d32.mean() <= 1.2 * d64.mean() + 2e-8
    assert d32.max() <= 1.5 * d64.max() + 1e-7


def split_chains(flat, lengths):
    out, at = [], 0
    for n in lengths:
        out.append(flat[at:at + n])
        at += n
    return out


def oracle_ragged(models, kws, sets, lengths, s0, steps):
    chains = split_chains(s0, lengths)
    res = [None] * len(lengths)
    for key in sorted(set(zip(sets, lengths))):
        idx = [i for i in range(len(lengths)) if (sets[i], lengths[i]) == key]
        st = key[0]
        out = O.move_real(ORACLE_MODEL[models[st]], np.stack([chains[i] for i in idx]), steps, 0.05, threads=8,
                          **kws[st])
        for j, i in enumerate(idx):
            res[i] = out[j]
    return np.concatenate(res)


def test_fp32c_ragged_toda_persistent(gpu):
    """The C4 layout at a reduced size: Toda chains of 32..2048 sites in one
    persistent launch, FP32C, against the oracle and the FP64 kernel."""
    N = gpu
    lengths = N.ic_ragged_lengths(400, [32, 64, 128, 256, 512, 1024, 2048], 4)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.25, 4)
    out = {}
    for prec in (N.FP64, N.FP32C):
        with N.Ensemble.ragged(N.TODA, lengths, 0.05, alpha=1.0, precision=prec) as ens:
            desc = ens.describe()
            ens.upload(s0)
            ens.move(100)
            out[prec] = ens.download()
    banner(desc)
    assert "ragged_persistent" in desc and "precision=fp32c" in desc
    lengths = list(lengths)
    want = oracle_ragged({0: 2}, {0: dict(alpha=1.0)}, [0] * len(lengths), lengths, s0, 100)
    err = chain_errors(out[N.FP32C], want, lengths)
    d64 = chain_errors(out[N.FP32C], out[N.FP64], lengths)
    banner(f"  100 steps: per-chain err against the oracle max {err.max():.2e}, against FP64 max {d64.max():.2e}")
    assert err.max() <= RAGGED_TOL and d64.max() <= RAGGED_TOL


def test_fp32c_ragged_short_chains_every_persistent_width(gpu):
    """Short FPUT chains (8..512 sites) in FP32C: the persistent kernel with
    8 sites per lane by default, then the widths that hold 8-site chains (8,
    and 16 at half width) and the per-class launches, which the autotuner
    times; runs before and after it checked against the oracle."""
    N = gpu
    lengths = N.ic_ragged_lengths(500, [8, 16, 32, 64, 128, 256, 512], 9)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.5, 9)
    want = oracle_ragged({0: 1}, {0: dict(alpha=0.25, beta=1.0)}, [0] * len(lengths), list(lengths), s0, 60)
    with N.Ensemble.ragged(N.FPUT, lengths, 0.05, alpha=0.25, beta=1.0, precision=N.FP32C) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(60)
        err0 = chain_errors(ens.download(), want, list(lengths)).max()
        report = ens.autotune(5)
        ens.upload(s0)
        ens.move(60)
        err1 = chain_errors(ens.download(), want, list(lengths)).max()
        desc2 = ens.describe()
    banner(f"{desc}: {err0:.2e}; autotune: {report}; {desc2}: {err1:.2e}")
    assert "ragged_persistent sites_per_lane=8" in desc and "precision=fp32c" in desc
    for w in ("S=8 ", "S=8/half", "S=16/half", "per-class"):
        assert w in report, report
    assert err0 <= RAGGED_TOL and err1 <= RAGGED_TOL


def test_fp32c_mixed_fput_toda(gpu):
    """FPUT and Toda chains of three lengths in one FP32C context: per-chain
    parity with the oracle, and the kernel the autotuner picks giving the same
    trajectory (another shape may take the other form of the Toda exponential,
    which differs in the last FP64 bits: bounded at FP32C rounding)."""
    N = gpu
    lengths = N.ic_ragged_lengths(300, [64, 256, 1024], 6)
    sets = list(np.arange(len(lengths)) % 2)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.25, 6)
    params = [dict(model=N.FPUT, beta=1.0, precision=N.FP32C), dict(model=N.TODA, alpha=1.0, precision=N.FP32C)]
    with N.Ensemble.mixed(params, lengths, sets, 0.05) as ens:
        desc = ens.describe()
        ens.upload(s0)
        ens.move(100)
        got = ens.download()
        report = ens.autotune(10)
        ens.upload(s0)
        ens.move(100)
        again = ens.download()
        desc2 = ens.describe()
    banner(f"{desc}; autotune: {report}; {desc2}")
    assert "mixed" in desc and "precision=fp32c" in desc
    want = oracle_ragged({0: 1, 1: 2}, {0: dict(beta=1.0), 1: dict(alpha=1.0)}, sets, list(lengths), s0, 100)
    err = chain_errors(got, want, list(lengths))
    banner(f"  per-chain err against the oracle max {err.max():.2e}; after autotune {chain_errors(again, got, list(lengths)).max():.2e}")
    assert err.max() <= RAGGED_TOL
    assert chain_errors(again, got, list(lengths)).max() <= 1e-7


def test_fp32c_batches_graph_and_autotune(gpu):
    """Two batches of 10 steps against one of 20 (a batch boundary stores the
    FP64 sum of each pair and reloads it, and the merged c7 + c0 drift of one
    batch runs as two FP32 increments: rounding-level differences), graph
    replays identical to the eager path, and the autotuner choosing among the
    FP32C shapes only."""
    N = gpu
    s0 = N.ic_plane_waves(256, 64, 0.0, 1, 32, 0.25, 8)
    with N.Ensemble(N.TODA, 256, 64, 0.05, alpha=1.0, precision=N.FP32C) as a, \
            N.Ensemble(N.TODA, 256, 64, 0.05, alpha=1.0, precision=N.FP32C) as b:
        a.upload(s0)
        a.move(10)
        a.move(10)
        b.upload(s0)
        b.move(20)
        d = chain_errors(a.download(), b.download()).max()
        banner(f"10 + 10 vs 20 steps: {d:.2e}")
        assert d <= 1e-7
        a.upload(s0)
        b.upload(s0)
        for _ in range(3):
            ea, sa = a.batch(10)
            b.move(10)
            eb, sb = b.linenergies()
            assert np.array_equal(a.download(), b.download())
            assert np.array_equal(ea, eb) and np.array_equal(sa, sb)
        report = a.autotune(5)
        banner(f"autotune: {report}; {a.describe()}")
        assert "split" not in report and "tiled" not in report and "precision=fp32c" in a.describe()


@pytest.mark.parametrize("model,n,kw", [(0, 64, dict(m=1.0, beta=1.0)), (1, 100, dict(beta=1.0)),
                                        (2, 4096, dict(alpha=1.0)), (3, 64, dict(beta=1.0))])
def test_fp32c_refusals(gpu, model, n, kw):
    """FP32C real models: FPUT and Toda, power-of-two chains up to 2048 sites."""
    N = gpu
    masses = np.ones(n) if model == N.DDNKG else None
    with pytest.raises(N.NlcError):
        N.Ensemble(model, n, 4, 0.05, masses=masses, precision=N.FP32C, **kw)
    if model == N.FPUT:
        with pytest.raises(N.NlcError):
            N.Ensemble(N.FPUT, 64, 4, 0.05, beta=1.0, precision=N.FP32C, kernel=N.KERNEL_SPLIT)
        with pytest.raises(N.NlcError, match="power-of-two"):
            N.Ensemble.ragged(N.FPUT, np.array([64, 100, 64], dtype=np.uint32), 0.05, beta=1.0, precision=N.FP32C)
        with pytest.raises(N.NlcError):  # one precision per mixed ensemble
            N.Ensemble.mixed([dict(model=N.FPUT, beta=1.0, precision=N.FP32C), dict(model=N.TODA, alpha=1.0)],
                             np.array([64, 64], dtype=np.uint32), [0, 1], 0.05)
