"""CPU checks of the benchmark workloads' definitions (no GPU).

The ragged Toda workload (alpha = 1, dt = 0.05, 65536 chains with lengths
uniform over {32, ..., 2048}, seed 4) is specified with per-chain energy
densities U[0.05, 0.5].  At that upper end the longest chains start with bond
stretches near 3 (e^{2 alpha x} ~ 400), where dt = 0.05 leaves the explicit
integrator's stability region: the FP64 oracle itself goes non-finite within
the first 1000-step batch.  bench.py therefore draws the densities from
U[0.05, 0.25]; these tests pin both facts on the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors, chain_rel_err, oracle_parallel

CHOICES = [32, 64, 128, 256, 512, 1024, 2048]
# chains (caller order) of the seed-4 ensemble at energy density up to 0.5 whose
# FP64 oracle run goes non-finite within the first 1000-step batch
DIVERGING = [4634, 9226, 19670, 39921, 44023]


@pytest.fixture(scope="module")
def ragged():
    import nlchains_amd as N
    lengths = N.ic_ragged_lengths(65536, CHOICES, 4)
    offs = np.concatenate([[0], np.cumsum(lengths.astype(np.int64))])
    half = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.5, 4)
    quarter = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.25, 4)
    return lengths, offs, half, quarter


def max_bond_stretch(chain):
    q = chain[:, 0]
    return np.abs(np.diff(np.concatenate([q, q[:1]]))).max()


def test_ragged_toda_half_density_diverges_in_the_oracle(ragged):
    lengths, offs, half, _ = ragged
    assert all(lengths[i] == 2048 for i in DIVERGING)
    chains = np.stack([half[offs[i]:offs[i + 1]] for i in DIVERGING])
    assert all(max_bond_stretch(c) > 3.0 for c in chains)
    out = oracle_parallel(O.move_real, chains, model=O.TODA, steps=1000, dt=0.05, alpha=1.0)
    assert not any(np.isfinite(o).all() for o in out)


def test_ragged_toda_bench_density_stays_finite(ragged):
    """At U[0.05, 0.25] the eight most stretched 2048-site chains (bond
    stretch at most about 2.3) run a whole 1000-step batch finitely and
    conserve energy to the integrator's truncation error."""
    lengths, offs, _, quarter = ragged
    big = np.nonzero(lengths == 2048)[0]
    stretch = np.array([max_bond_stretch(quarter[offs[i]:offs[i + 1]]) for i in big])
    top = big[np.argsort(stretch)[-8:]]
    assert stretch.max() < 2.5
    chains = np.stack([quarter[offs[i]:offs[i + 1]] for i in top])
    out = oracle_parallel(O.move_real, chains, model=O.TODA, steps=1000, dt=0.05, alpha=1.0)
    assert np.isfinite(out).all()
    E0 = O.chain_energy(O.TODA, chains, alpha=1.0)
    E1 = O.chain_energy(O.TODA, out, alpha=1.0)
    assert np.abs(E1 / E0 - 1).max() < 1e-3


def test_chain_errors_catch_a_wrong_small_chain():
    """A planted relative error of 1e-8 in a chain whose amplitude is 1e-4 of
    the largest chain's: hidden by a shard-wide max-norm, caught chain by
    chain."""
    rng = np.random.default_rng(0)
    want = rng.normal(size=(16, 64, 2))
    want[5] *= 1e-4
    got = want.copy()
    got[5, 7, 1] += 1e-8 * np.abs(want[5]).max()
    shard_wide = np.abs(got - want).max() / np.abs(want).max()
    assert shard_wide < 1e-11
    assert chain_rel_err(got, want) == pytest.approx(1e-8, rel=1e-6)
    err = chain_errors(got, want)
    assert np.argmax(err) == 5 and (np.delete(err, 5) == 0).all()
    # ragged layout: the same chains concatenated with their lengths
    assert chain_rel_err(got.reshape(-1, 2), want.reshape(-1, 2), [64] * 16) == pytest.approx(1e-8, rel=1e-6)
    # a NaN fails its chain
    got[2, 0, 0] = np.nan
    assert np.isinf(chain_errors(got, want)[2])
