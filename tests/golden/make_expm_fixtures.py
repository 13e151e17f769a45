"""Writes expm_fixtures.npz: matrices and their exponentials computed with
scipy.linalg.expm (Al-Mohy & Higham 2009), used to pin the oracle's own
Pade exponential and the Krylov propagator builders. Run once in the build
container: python tests/golden/make_expm_fixtures.py"""
import os

import numpy as np
import scipy
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))


def energy_conserving(n, rng, density=0.3):
    """B = K - diag(colsum K): columns sum to zero, off-diagonals >= 0."""
    K = rng.random((n, n)) * (rng.random((n, n)) < density)
    np.fill_diagonal(K, 0.0)
    return K - np.diag(K.sum(0))


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    cases = []
    # general dense matrices of several norms (exercise scaling and squaring)
    for n, scale in ((3, 0.01), (5, 0.3), (8, 1.0), (16, 4.0), (32, 20.0)):
        cases.append((f"dense_n{n}", rng.standard_normal((n, n)) * scale / np.sqrt(n)))
    # energy-conserving generators like t*B of a collision operator
    for n, t in ((10, 0.5), (64, 2.0), (128, 5.0)):
        cases.append((f"ecop_n{n}", t * energy_conserving(n, rng)))
    for name, A in cases:
        out[name + "_A"] = A
        out[name + "_E"] = scipy.linalg.expm(A)
    out["scipy_version"] = np.array(scipy.__version__)
    np.savez_compressed(os.path.join(HERE, "expm_fixtures.npz"), **out)
    print("wrote", sorted(k for k in out if k.endswith("_A")))


if __name__ == "__main__":
    main()
