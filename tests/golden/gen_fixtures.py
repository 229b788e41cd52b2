"""Writes tests/golden/fixtures.npz: DFT vectors from numpy's pocketfft and an
eigensystem from LAPACK (numpy.linalg.eigh) for the dDNKG operator.

The reference holds no golden vectors for its hot path; these pin the oracle's
restatements of FFTW3 (orc_fft) and Armadillo eig_sym (orc_eigensystem) against
independent library implementations.  Regenerate with
    python tests/golden/gen_fixtures.py
"""
from pathlib import Path

import numpy as np


def main():
    rng = np.random.default_rng(2024)
    out = {}
    for n in (7, 16, 60, 256):
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        out[f"fft_in_{n}"] = z
        out[f"fft_out_{n}"] = np.fft.fft(z)
    n = 48
    mp2 = 2 + rng.uniform(0.5, 1.5, size=n)
    A = np.diag(mp2) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    lam, V = np.linalg.eigh(A)
    out["eig_mp2"] = mp2
    out["eig_omega"] = np.sqrt(lam)
    out["eig_V"] = V
    np.savez(Path(__file__).with_name("fixtures.npz"), **out)


if __name__ == "__main__":
    main()
