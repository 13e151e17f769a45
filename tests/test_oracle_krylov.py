"""Oracle Krylov propagator pinned by scipy golden vectors, closed forms and
the Arnoldi relations of both projection modes."""
import os

import numpy as np
import pytest

import oracle_lib as orc

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "expm_fixtures.npz"))
CASES = sorted(k[:-2] for k in FIX.files if k.endswith("_A"))


@pytest.mark.parametrize("name", CASES)
def test_expm_matches_scipy_golden(name):
    A, E = FIX[name + "_A"], FIX[name + "_E"]
    got = orc.expm(A)
    scale = np.abs(E).max()
    assert np.abs(got - E).max() <= 1e-12 * max(scale, 1.0) * (1 + np.abs(A).sum(0).max())


def test_expm_closed_forms():
    # diagonal
    d = np.array([-3.0, -0.5, 0.0, 0.7])
    np.testing.assert_allclose(orc.expm(np.diag(d)), np.diag(np.exp(d)), rtol=1e-14)
    # 2x2 energy-conserving B = [[-a, b], [a, -b]]: exp(tB) = I + (1-e^{-(a+b)t})/(a+b) B
    a, b, t = 0.7, 0.2, 3.0
    B = np.array([[-a, b], [a, -b]])
    ref = np.eye(2) + (1 - np.exp(-(a + b) * t)) / (a + b) * B
    np.testing.assert_allclose(orc.expm(t * B), ref, rtol=1e-13, atol=1e-15)
    # nilpotent: exp(N) = I + N + N^2/2
    N = np.array([[0.0, 2.0, 3.0], [0.0, 0.0, 5.0], [0.0, 0.0, 0.0]])
    np.testing.assert_allclose(orc.expm(N), np.eye(3) + N + N @ N / 2, rtol=1e-14)


@pytest.mark.parametrize("mode", [0, 1])
def test_arnoldi_relation(mode):
    """t A V_k = V H_k + h e^T for the columns before the last (both the
    reference loop over j<k and the textbook loop over j<=k)."""
    rng = np.random.default_rng(mode)
    A = rng.standard_normal((40, 40)) / 6
    b = rng.standard_normal(40)
    n, t = 12, 0.8
    V, H = orc.arnoldi(A, b, n, t, mode)
    lhs = t * A @ V[:, : n - 1]
    rhs = V @ H[:, : n - 1]
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)
    np.testing.assert_allclose(V[:, 0], b / np.linalg.norm(b), rtol=1e-14)
    if mode == 1:
        np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)
    else:
        # the reference never forms H(k,k) (collision_operator.cpp:833)
        assert np.all(np.diag(H) == 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_build_P_exact_when_krylov_space_is_full(mode):
    """m = N spans the whole space: the Krylov polynomial equals exp(tA) b."""
    A = FIX["ecop_n10_A"]
    P = orc.build_P(A, 1.0, 10, mode)
    # the reference loop's basis is not orthogonal and loses ~7 digits here
    np.testing.assert_allclose(P, FIX["ecop_n10_E"], atol=1e-9 if mode == 1 else 1e-6)


def test_build_P_converges_on_golden_operator():
    A = FIX["ecop_n64_A"]
    E = FIX["ecop_n64_E"]
    errs = [np.abs(orc.build_P(A, 1.0, m, 1) - E).max() for m in (6, 12, 24)]
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-10


def test_krylov_size_rule():
    """ksize = smallest n >= 2 with error_bound(n)/N^2 < eps (collision_operator.cpp:1021-1029)."""
    A = FIX["ecop_n10_A"] * 0.1
    k = orc.krylov_size(A, 1.0, 1e-8)
    assert k >= 2
    assert orc.error_bound(A, k, 1.0) / 100 < 1e-8
    if k > 2:
        assert orc.error_bound(A, k - 1, 1.0) / 100 >= 1e-8


def test_project_columns_makes_columns_sum_zero():
    rng = np.random.default_rng(5)
    B = rng.standard_normal((30, 30))
    Bp = orc.project_columns(B)
    assert np.abs(Bp.sum(0)).max() < 1e-14
    # the correction is a constant per column
    d = Bp - B
    np.testing.assert_allclose(d, np.broadcast_to(d[0], d.shape), atol=1e-15)


def test_synthetic_operator_invariants():
    m = orc.synthetic_modes(500, 2)
    T = 300.0
    C = orc.bose_einstein_kernel(m["omega"], T)
    B = orc.synthetic_B(C, m["w"], 0.05, 7)
    off = B - np.diag(np.diag(B))
    assert off.min() >= 0
    assert np.abs(B.sum(0)).max() < 1e-12 * np.abs(np.diag(B)).max()     # 1^T B = 0
    assert np.abs(B @ C).max() < 1e-12 * np.abs(np.diag(B) * C).max()     # B C = 0
    S = B * C[None, :]                                                     # B diag(C) symmetric
    np.testing.assert_allclose(S, S.T, rtol=1e-12, atol=1e-30)
    dens = (off > 0).mean()
    assert 0.04 < dens < 0.06
    # P = exp(tB) is column-stochastic and nonnegative, with P C = C
    E = orc.expm(5.0 * B)
    assert E.min() > -1e-12
    np.testing.assert_allclose(E.sum(0), 1.0, atol=1e-12)
    np.testing.assert_allclose(E @ C, C, rtol=1e-10)
