"""The batched form: ragged ensembles in one persistent launch, hipGraph
batches and the autotuner, all checked against the oracle (FP64)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_errors, chain_rel_err

pytestmark = pytest.mark.gpu

ORACLE_MODEL = {0: O.DNKG, 1: O.FPUT, 2: O.TODA}


def banner(msg):
    print(f"[gpu-batched] {msg}", flush=True)
    sys.stdout.flush()


def split_chains(flat, lengths):
    out, at = [], 0
    for n in lengths:
        out.append(flat[at:at + n])
        at += n
    return out


def oracle_ragged(model, lengths, s0, steps, dt, kw):
    """Advance every chain with the oracle, grouped by length."""
    chains = split_chains(s0, lengths)
    res = [None] * len(lengths)
    for n in sorted(set(lengths)):
        idx = [i for i, l in enumerate(lengths) if l == n]
        out = O.move_real(ORACLE_MODEL[model], np.stack([chains[i] for i in idx]), steps, dt, **kw)
        for j, i in enumerate(idx):
            res[i] = out[j]
    return np.concatenate(res)


@pytest.mark.parametrize("model,kw,choices", [
    (2, dict(alpha=1.0), [32, 64, 128, 256, 512]),
    (0, dict(m=1.0, beta=1.0), [32, 128, 1024, 2048]),
    (1, dict(alpha=0.25, beta=1.0), [64, 256]),
    (0, dict(m=1.0, beta=1.0), [8, 16, 32, 64]),
    (2, dict(alpha=1.0), [70, 100, 300]),  # no persistent shape: one uniform kernel per class
    (1, dict(beta=1.0), [5, 24, 45, 128]),  # short classes: thread per chain, 4 lanes
    (0, dict(m=1.0, beta=0.3), [32, 128, 1024]),  # scaled cubic form, sigma = sqrt(0.3)
    (1, dict(alpha=0.2), [64, 256]),  # scaled quadratic form, sigma = 0.2
])
def test_ragged_parity(gpu, model, kw, choices):
    N = gpu
    lengths = N.ic_ragged_lengths(301, choices, 4)
    s0 = N.ic_ragged_plane_waves(lengths, kw.get("m", 0.0), 0.05, 0.05, 0.5, 4)
    with N.Ensemble.ragged(model, lengths, 0.05, **kw) as ens:
        desc = ens.describe()
        banner(desc)
        ens.upload(s0)
        assert np.array_equal(ens.download(), s0)  # upload/download permutation round trip
        ens.move(100)
        got = ens.download()
        E, _ = ens.chain_energies()
    want = oracle_ragged(model, list(lengths), s0, 100, 0.05, kw)
    err = chain_rel_err(got, want, lengths)
    banner(f"  largest per-chain rel err {err:.3e}")
    assert err <= 1e-10
    Ew = np.concatenate([O.chain_energy(ORACLE_MODEL[model], c[None], **kw) for c in split_chains(want, lengths)])
    np.testing.assert_allclose(E, Ew, rtol=1e-11)


def test_ragged_linenergies_per_class(gpu):
    N = gpu
    lengths = N.ic_ragged_lengths(200, [32, 64, 256], 7)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.5, 7)
    s0 = s0 + 0.01 * np.random.default_rng(0).normal(size=s0.shape)
    with N.Ensemble.ragged(N.TODA, lengths, 0.05, alpha=1.0) as ens:
        ens.upload(s0)
        e_gpu, s_gpu = ens.linenergies()
        cl = list(ens.class_lengths)
    chains = split_chains(s0, lengths)
    at = 0
    for ci, n in enumerate(cl):
        group = np.stack([c for c, l in zip(chains, lengths) if l == n])
        e = O.calc_linenergies(O.linenergies_real(group, O.dispersion_real(n, 0.0)), 0.5 / len(group) / n, True)
        wt, inf = O.entropies(e, True)
        np.testing.assert_allclose(e_gpu[at:at + n], e, rtol=0, atol=1e-9 * e.max())
        assert abs(s_gpu[2 * ci] - wt) <= 1e-9 * abs(wt) and abs(s_gpu[2 * ci + 1] - inf) <= 1e-9 * abs(inf)
        at += n


@pytest.mark.parametrize("kind", ["wave", "ragged", "tiled"])
def test_graph_batches_match_eager(gpu, kind):
    """nlc_batch replays a captured graph; every replay (with eager work and
    synchronisations in between) equals the eager move + energy pass.  The
    tiled case (20000 sites, 40 steps = launches of 18, 18 and 4) ends its
    batch with a device-to-device copy node."""
    N = gpu
    steps = 25
    if kind == "ragged":
        lengths = N.ic_ragged_lengths(120, [32, 64, 128], 2)
        s0 = N.ic_ragged_plane_waves(lengths, 1.0, 0.05, 0.1, 0.3, 2)
        mk = lambda: N.Ensemble.ragged(N.DNKG, lengths, 0.05, m=1.0, beta=1.0)  # noqa: E731
    elif kind == "tiled":
        s0 = N.ic_plane_waves(20000, 3, 1.0, 0, 64, 0.3, 2)
        mk = lambda: N.Ensemble(N.DNKG, 20000, 3, 0.05, m=1.0, beta=1.0)  # noqa: E731
        steps = 40
    else:
        s0 = N.ic_plane_waves(1024, 64, 1.0, 0, 128, 0.5, 2)
        mk = lambda: N.Ensemble(N.DNKG, 1024, 64, 0.05, m=1.0, beta=1.0)  # noqa: E731
    with mk() as g, mk() as e:
        if kind == "tiled":
            assert "kernel=tiled" in g.describe()
        g.upload(s0)
        e.upload(s0)
        for rep in range(4):
            eg, sg = g.batch(steps)
            e.move(steps)
            ee, se = e.linenergies()
            assert np.array_equal(g.download(), e.download()), rep
            assert np.array_equal(eg, ee) and np.array_equal(sg, se), rep
            e.chain_energies()  # unrelated eager work between replays


@pytest.mark.parametrize("direct", [False, True])
@pytest.mark.parametrize("ragged", [False, True])
def test_queued_batches_match_synchronous(gpu, ragged, direct):
    """nlc_batch_submit / nlc_batch_collect, as graph replays and as direct
    launches: results come back in submission order, equal to nlc_batch's; the
    slot ring wraps; the state after the queue is the state after all queued
    batches; the queue depth is enforced; the mode cannot change while batches
    are queued."""
    N = gpu
    if ragged:
        lengths = N.ic_ragged_lengths(100, [32, 64, 256], 3)
        s0 = N.ic_ragged_plane_waves(lengths, 1.0, 0.05, 0.1, 0.3, 3)
        mk = lambda: N.Ensemble.ragged(N.FPUT, lengths, 0.05, beta=1.0)  # noqa: E731
    else:
        s0 = N.ic_plane_waves(64, 1024, 0.0, 1, 5, 0.1, 1)
        mk = lambda: N.Ensemble(N.FPUT, 64, 1024, 0.1, beta=1.0)  # noqa: E731
    with mk() as q, mk() as s:
        q.set_batch_mode(direct)
        q.upload(s0)
        s.upload(s0)
        want = [s.batch(20) for _ in range(10)]
        got = []
        for i in range(N.BATCH_SLOTS):
            q.batch_submit(20)
        with pytest.raises(N.NlcError):
            q.batch_submit(20)
        with pytest.raises(N.NlcError) as ex:
            q.set_batch_mode(not direct)
        assert ex.value.code == N.NLC_ESTATE
        got.append(q.batch_collect())
        for i in range(6):  # keep three queued while collecting, wrapping the ring
            q.batch_submit(20)
            got.append(q.batch_collect())
        while len(got) < 10:
            got.append(q.batch_collect())
        with pytest.raises(N.NlcError):
            q.batch_collect()
        for i, ((eg, sg), (ew, sw)) in enumerate(zip(got, want)):
            assert np.array_equal(eg, ew) and np.array_equal(sg, sw), i
        assert np.array_equal(q.download(), s.download())
        # queued work followed by eager calls: they see the state after the queue
        q.batch_submit(20)
        q.batch_submit(20)
        first = s.batch(20)
        second = s.batch(20)
        assert np.array_equal(q.download(), s.download())
        eq, sq = q.linenergies()
        assert np.array_equal(eq, second[0]) and np.array_equal(sq, second[1])
        for want_e, want_s in (first, second):
            eg, sg = q.batch_collect()
            assert np.array_equal(eg, want_e) and np.array_equal(sg, want_s)


def test_autotune_keeps_state_and_parity(gpu):
    N = gpu
    s0 = N.ic_plane_waves(64, 500, 0.0, 1, 8, 0.3, 5)
    with N.Ensemble(N.FPUT, 64, 500, 0.05, alpha=0.25, beta=1.0) as ens:
        ens.upload(s0)
        report = ens.autotune(10)
        banner(f"autotune: {report} -> {ens.describe()}")
        assert "*" in report and "queued batches: graph" in report
        assert np.array_equal(ens.download(), s0)
        ens.move(50)
        got = ens.download()
    want = O.move_real(O.FPUT, s0, 50, 0.05, alpha=0.25, beta=1.0)
    assert chain_rel_err(got, want) <= 1e-10


def test_autotune_ragged(gpu):
    N = gpu
    lengths = N.ic_ragged_lengths(400, [64, 128, 256], 9)
    s0 = N.ic_ragged_plane_waves(lengths, 1.0, 0.05, 0.1, 0.3, 9)
    with N.Ensemble.ragged(N.DNKG, lengths, 0.05, m=1.0, beta=1.0) as ens:
        ens.upload(s0)
        report = ens.autotune(5)
        banner(f"autotune ragged: {report}")
        assert all(f"S={s} " in report for s in (8, 16, 32))  # S <= n_min = 64 and 64 S >= n_max = 256
        assert np.array_equal(ens.download(), s0)
        ens.move(20)
        got = ens.download()
    want = oracle_ragged(0, list(lengths), s0, 20, 0.05, dict(m=1.0, beta=1.0))
    assert chain_rel_err(got, want, lengths) <= 1e-10


def test_ragged_toda_full_size_sampled(gpu):
    """The ragged Toda workload: 65536 chains, lengths uniform over
    {32, ..., 2048}; 32 chains of each length re-run on the oracle for 100
    steps, compared chain by chain."""
    N = gpu
    choices = [32, 64, 128, 256, 512, 1024, 2048]
    lengths = N.ic_ragged_lengths(65536, choices, 4)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.5, 4)
    steps = 100
    with N.Ensemble.ragged(N.TODA, lengths, 0.05, alpha=1.0) as ens:
        banner(ens.describe())
        ens.upload(s0)
        e, s = ens.batch(steps)
        got = ens.download()
    assert np.isfinite(e).all() and np.isfinite(s).all()
    offs = np.concatenate([[0], np.cumsum(lengths.astype(np.int64))])
    rng = np.random.default_rng(1)
    for n in choices:
        idx = rng.choice(np.nonzero(lengths == n)[0], 32, replace=False)
        chains = np.stack([s0[offs[i]:offs[i + 1]] for i in idx])
        want = O.move_real(O.TODA, chains, steps, 0.05, alpha=1.0)
        g = np.stack([got[offs[i]:offs[i + 1]] for i in idx])
        err = chain_errors(g, want)
        banner(f"  n={n}: largest per-chain rel err {err.max():.3e}")
        assert err.max() <= 1e-10, n


MIXED_SETS = [dict(model=0, m=1.0, beta=1.0), dict(model=1, alpha=0.0, beta=1.0), dict(model=2, alpha=1.0)]


def oracle_mixed(sets, lengths, chain_set, s0, steps, dt):
    """Every chain advanced with the oracle of its own model and parameters."""
    chains = split_chains(s0, lengths)
    res = [None] * len(lengths)
    for si, kw in enumerate(sets):
        kw = {k: v for k, v in kw.items() if k != "model"}
        for n in sorted(set(lengths)):
            idx = [i for i in range(len(lengths)) if lengths[i] == n and chain_set[i] == si]
            if not idx:
                continue
            out = O.move_real(ORACLE_MODEL[sets[si]["model"]], np.stack([chains[i] for i in idx]), steps, dt, **kw)
            for j, i in enumerate(idx):
                res[i] = out[j]
    return np.concatenate(res)


@pytest.mark.parametrize("choices", [[32, 64, 128, 256, 512, 1024, 2048], [8, 16, 64, 100]])
def test_mixed_models_parity(gpu, choices):
    """DNKG, FPUT and Toda chains of several lengths in one context: one
    persistent launch per move for powers of two (per-class launches for
    lengths without a persistent shape), every chain against the oracle of its
    own model after 2 x 50 steps (the second launch only advances the state if
    the first reset the work queue), chain energies, and energies and
    entropies per (model, length) class."""
    N = gpu
    nchains = 420
    lengths = N.ic_ragged_lengths(nchains, choices, 11)
    chain_set = np.random.default_rng(12).integers(0, 3, nchains).astype(np.uint32)
    # DNKG chains need m > 0 for mode 0 (the IC helper takes the dispersion's m)
    s0 = np.concatenate([N.ic_ragged_plane_waves([n], 1.0 if cs == 0 else 0.0, 0.05, 0.05, 0.3, 100 + i)
                         for i, (n, cs) in enumerate(zip(lengths, chain_set))])
    with N.Ensemble.mixed(MIXED_SETS, lengths, chain_set, 0.05) as ens:
        desc = ens.describe()
        banner(desc)
        assert "mixed" in desc and "DNKG+FPUT+Toda" in desc
        assert ("ragged_persistent" in desc) == (100 not in choices)
        ens.upload(s0)
        assert np.array_equal(ens.download(), s0)
        ens.move(50)
        ens.move(50)
        got = ens.download()
        E, _ = ens.chain_energies()
        e_gpu, s_gpu = ens.linenergies()
        cl, cset = list(ens.class_lengths), list(ens.class_sets)
    want = oracle_mixed(MIXED_SETS, list(lengths), chain_set, s0, 100, 0.05)
    err = chain_errors(got, want, lengths)
    banner(f"  largest per-chain rel err {err.max():.3e}")
    assert err.max() <= 1e-10
    wchains = split_chains(want, lengths)
    for i in range(nchains):
        kw = {k: v for k, v in MIXED_SETS[chain_set[i]].items() if k != "model"}
        Ew = O.chain_energy(ORACLE_MODEL[MIXED_SETS[chain_set[i]]["model"]], wchains[i][None], **kw)[0]
        assert abs(E[i] - Ew) <= 1e-11 * abs(Ew), i
    # classes in ascending (set, length); each with its model's dispersion and a0is0
    assert cset == sorted(cset) and len(cl) == len(set(zip(cset, cl)))
    at = 0
    for ci, (n, si) in enumerate(zip(cl, cset)):
        group = np.stack([wchains[i] for i in range(nchains) if lengths[i] == n and chain_set[i] == si])
        m = MIXED_SETS[si].get("m", 0.0)
        e = O.calc_linenergies(O.linenergies_real(group, O.dispersion_real(n, m)), 0.5 / len(group) / n, m == 0.0)
        np.testing.assert_allclose(e_gpu[at:at + n], e, rtol=0, atol=1e-9 * e.max())
        at += n


def test_persistent_variants_and_queue_reset(gpu):
    """Every persistent variant the autotuner offers (S sites per lane, with and
    without the S/2 items) gives the oracle's trajectory, also over several
    launches and graph replays (the work queue is reset by each launch)."""
    N = gpu
    lengths = N.ic_ragged_lengths(600, [32, 64, 128, 256, 512, 1024, 2048], 21)
    s0 = N.ic_ragged_plane_waves(lengths, 0.0, 0.05, 0.05, 0.25, 21)
    want = oracle_ragged(2, list(lengths), s0, 60, 0.05, dict(alpha=1.0))
    with N.Ensemble.ragged(N.TODA, lengths, 0.05, alpha=1.0) as ens:
        ens.upload(s0)
        report = ens.autotune(10)
        banner(f"autotune: {report} -> {ens.describe()}")
        assert "S=32 " in report and "S=32/half" in report and "per-class" in report
        ens.move(20)
        ens.batch(20)
        ens.batch(20)
        got = ens.download()
    assert chain_rel_err(got, want, lengths) <= 1e-10


def test_mixed_rejects_bad_sets(gpu):
    N = gpu
    lengths = np.array([64, 128], np.uint32)
    for sets, cs in (([dict(model=3, beta=1.0)], [0, 0]), ([dict(model=4, beta=1.0)], [0, 0]),
                     (MIXED_SETS, [0, 3]), (MIXED_SETS + MIXED_SETS[:1], [0, 1])):
        with pytest.raises(N.NlcError):
            N.Ensemble.mixed(sets, lengths, cs, 0.05)
