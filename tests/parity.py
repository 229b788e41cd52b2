"""Error measures shared by the parity tests.

The acceptance bar for states ("q/p within 1e-10 relative after 100 steps") is
a statement about each trajectory, so states are compared chain by chain: the
error of a chain is ||got - want||_inf / ||want||_inf over that chain's sites,
and a test reports the largest over all chains.  A wrong low-amplitude chain
therefore cannot hide under the largest chain's maximum.  Energy vectors
(one value per mode) are compared relative to their largest entry.
"""
import numpy as np


def chain_errors(got, want, lengths=None):
    """Per-chain relative max-norm errors.  Uniform ensembles: arrays whose
    first axis is the chain (e.g. (copies, n, 2)); ragged ensembles: the
    chains concatenated, (sites, 2), with their lengths in order."""
    got, want = _real(got), _real(want)
    assert got.shape == want.shape, (got.shape, want.shape)
    if lengths is None:
        g = got.reshape(got.shape[0], -1)
        w = want.reshape(want.shape[0], -1)
        num = np.abs(g - w).max(axis=1)
        den = np.abs(w).max(axis=1)
    else:
        lengths = np.asarray(lengths, dtype=np.int64)
        assert lengths.sum() == got.shape[0]
        starts = np.concatenate([[0], np.cumsum(lengths)[:-1]])
        d = np.abs(got - want).reshape(got.shape[0], -1).max(axis=1)
        a = np.abs(want).reshape(want.shape[0], -1).max(axis=1)
        num = np.maximum.reduceat(d, starts)
        den = np.maximum.reduceat(a, starts)
    with np.errstate(divide="ignore", invalid="ignore"):
        err = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    # a chain with a non-finite value on either side fails
    return np.where(np.isfinite(num) & np.isfinite(den), err, np.inf)


def _real(a):
    """Complex arrays as (..., 2) real pairs; everything as float64."""
    a = np.asarray(a)
    if np.iscomplexobj(a):
        a = np.stack([a.real, a.imag], axis=-1)
    return np.asarray(a, dtype=np.float64)


def chain_rel_err(got, want, lengths=None):
    """The largest per-chain relative max-norm error (see chain_errors)."""
    return float(chain_errors(got, want, lengths).max())


def vec_rel_err(got, want):
    """||got - want||_inf / ||want||_inf of a vector (energies per mode)."""
    got, want = _real(got), _real(want)
    return float(np.abs(got - want).max() / np.abs(want).max())


def oracle_parallel(fn, state, workers=8, **kw):
    """fn(state=chunk, **kw) on chunks of chains in parallel threads (the
    oracle's ctypes calls release the GIL); the results concatenated."""
    from concurrent.futures import ThreadPoolExecutor
    state = np.asarray(state)
    chunks = [c for c in np.array_split(state, min(workers, len(state))) if len(c)]
    with ThreadPoolExecutor(len(chunks)) as ex:
        parts = list(ex.map(lambda c: fn(state=c, **kw), chunks))
    return np.concatenate(parts)
