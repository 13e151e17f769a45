"""Negative control for the BF16 Krylov tolerance: the GPU build stores
dt*B and the Arnoldi basis as BF16 hi + lo pairs and forms each product as
A_hi V_hi + A_hi V_lo + A_lo V_hi with FP32 accumulation. A numpy emulation
of that arithmetic on the golden 128-mode operator (t = 1, m = 40, textbook
projections, the additive column-sum projection) lands within the BF16
tolerance the GPU tests use, and the same emulation with every lo part
dropped misses it by far more than a factor of ten: a build that silently
lost its compensation terms would fail those tests."""
import os

import numpy as np
import scipy.linalg

# the BF16 column-L1 tolerance of the GPU Krylov tests at N = 128 and 256
BF16_TOL = 5e-5


def _bf16(x):
    """round-to-nearest-even of float32 values to bfloat16, as float64"""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


def _split(x, keep_lo):
    f = np.asarray(x, np.float32).astype(np.float64)
    h = _bf16(f)
    return h, (_bf16(f - h) if keep_lo else np.zeros_like(h))


def _build(A, t, m, keep_lo):
    N = A.shape[0]
    Ah, Al = _split(t * A, keep_lo)
    P = np.zeros((N, N))
    for j in range(N):
        V = np.zeros((N, m + 1))
        H = np.zeros((m + 1, m))
        V[j, 0] = 1.0
        for k in range(m):
            vh, vl = _split(V[:, k], keep_lo)
            z = (Ah @ vh + Ah @ vl + Al @ vh).astype(np.float32).astype(np.float64)
            for i in range(k + 1):
                H[i, k] = V[:, i] @ z
                z = z - H[i, k] * V[:, i]
            H[k + 1, k] = np.linalg.norm(z)
            h, l = _split(z / H[k + 1, k], keep_lo)
            V[:, k + 1] = h + l
        P[:, j] = V[:, :m] @ scipy.linalg.expm(H[:m, :m])[:, 0]
    return P + (1.0 - P.sum(0)) / N


def test_bf16_tolerance_separates_hi_lo_from_hi_only():
    fix = np.load(os.path.join(os.path.dirname(__file__), "golden", "expm_fixtures.npz"))
    A, E = fix["ecop_n128_A"], fix["ecop_n128_E"]
    l1 = {lo: float(np.abs(_build(A, 1.0, 40, lo) - E).sum(0).max()) for lo in (True, False)}
    assert l1[True] < BF16_TOL, l1
    assert l1[False] > 10 * BF16_TOL, l1
