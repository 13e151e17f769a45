"""Oracle physics helpers pinned by closed forms and independent routines."""
import numpy as np
import pytest
import scipy.interpolate

import oracle_lib as orc

KB = 1.3806488e-23
HBAR = 1.054571726e-34


def test_bose_einstein_closed_form():
    om = np.array([0.1, 1.0, 10.0, 50.0, 100.0])
    for T in (10.0, 300.0, 1000.0):
        x = 1e12 * HBAR * om / (KB * T)
        np.testing.assert_allclose(orc.bose_einstein(om, T), 1.0 / np.expm1(x), rtol=1e-13)


def test_kernel_identity_and_cutoffs():
    """(x/sinh x)^2 with x = hw/2kT equals (hw/kT)^2 n (n+1) (the derivative
    of the Bose-Einstein occupation), a different formula for the same value."""
    om = np.linspace(0.05, 90.0, 200)
    for T in (50.0, 300.0):
        n = orc.bose_einstein(om, T)
        y = 1e12 * HBAR * om / (KB * T)
        np.testing.assert_allclose(orc.bose_einstein_kernel(om, T), y * y * n * (n + 1), rtol=1e-11)
    # cutoffs of utilities.hpp:401-415: x > 300 -> 0, x == 0 -> 1
    assert orc.bose_einstein_kernel(np.array([1e6]), 1.0)[0] == 0.0
    assert orc.bose_einstein_kernel(np.array([0.0]), 300.0)[0] == 1.0


def test_calc_cv_sums_modes():
    om = np.linspace(1.0, 80.0, 50)
    V = 3.7
    cv = orc.calc_cv(om, V, 300.0, nq=2)
    assert cv == pytest.approx(KB * orc.bose_einstein_kernel(om, 300.0).sum() / 2 / V, rel=1e-14)
    # classical limit: every mode contributes kB
    assert orc.calc_cv(np.full(10, 1e-3), 1.0, 1000.0) == pytest.approx(10 * KB, rel=1e-9)


def test_wall_flux_linear_response_matches_calc_dotEeff_wall():
    """For a small wall offset dT the isothermal wall flux (sampling.cpp:
    266-272) must equal the reference's separate linear formula
    calc_dotEeff_wall = kB |dT| sum_{n.v>=0} kernel * n.v / nq (RTA_MC2D.cpp:
    231-259) per unit-cell volume."""
    m = orc.synthetic_modes(400, 9)
    vuc = 2.5
    for nx, ny in ((1.0, 0.0), (0.0, -1.0)):
        dT = 1e-4
        _, _, J = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.0 + dT, 300.0, nx, ny, vuc)
        vn = nx * m["vx"] + ny * m["vy"]
        lin = KB * dT * (orc.bose_einstein_kernel(m["omega"], 300.0) * np.where(vn > 0, vn, 0)).sum() / vuc
        assert J == pytest.approx(lin, rel=2e-4)
        _, _, Jm = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.0 - dT, 300.0, nx, ny, vuc)
        assert Jm < 0 and Jm == pytest.approx(-lin, rel=2e-4)


def test_wall_distribution_only_emits_inward():
    m = orc.synthetic_modes(300, 4)
    _, cum, _ = orc.wall_dist(m["omega"], m["vx"], m["vy"], 310.0, 300.0, 0.0, 1.0, 1.0)
    p = np.diff(np.concatenate([[0.0], cum]))
    assert (p[m["vy"] <= 0] == 0).all() and (p[m["vy"] > 0] > 0).all()
    assert cum[-1] == pytest.approx(1.0)


def test_spline_matches_scipy_natural():
    X = np.linspace(10, 1000, 100)
    Y = np.tanh((X - 300) / 150.0) * 1e-3 + 1e-6 * X
    q = np.concatenate([np.linspace(5, 1010, 300), X])
    ref = scipy.interpolate.CubicSpline(X, Y, bc_type="natural", extrapolate=True)(q)
    np.testing.assert_allclose(orc.spline(X, Y, q), ref, rtol=1e-9, atol=1e-15)


def test_grid_sampling_brute_force_and_chi2():
    p = np.array([0.5, 0.0, 2.0, 1.0, 0.5])
    s = orc.grid_sample(p, 1, 2, 200000)
    assert (s != 1).all()
    counts = np.bincount(s, minlength=5)
    expect = p / p.sum() * len(s)
    nz = expect > 0
    chi2 = ((counts[nz] - expect[nz]) ** 2 / expect[nz]).sum()
    assert chi2 < 25  # 3 dof


def test_propagator_H_enumeration():
    """propagator_H::get (collision_operator.cpp:854-904): lower_bound of R on
    |signed cumulative|, sign of the chosen entry."""
    b = np.array([0.2, -0.1, 0.0, 0.7])
    R = np.array([0.0, 0.1, 0.2, 0.25, 0.3, 0.31, 0.9999, 1.0])
    sg, idx = orc.propH_get(b, R)
    assert list(idx) == [0, 0, 0, 1, 1, 3, 3, 3]
    assert list(sg) == [1, 1, 1, -1, -1, 1, 1, 1]


def test_propagator_sampling_mean_is_column():
    """E[sign] per mode under the re-queue rule equals P(:,j)/sum_i P_ij:
    sampling with probability |P|/S and requeueing twice on a sign flip."""
    rng = np.random.default_rng(0)
    b = np.array([0.6, -0.15, 0.35, 0.2])   # column sum 1, |b| sum 1.3
    R = rng.random(400000)
    sg, idx = orc.propH_get(b, R)
    # each draw has weight: expected children multiply by 1/(1 - 2 p_neg)
    pneg = 0.15 / 1.3
    est = np.bincount(idx, weights=sg, minlength=4) / len(R) / (1 - 2 * pneg)
    np.testing.assert_allclose(est, b, atol=5e-3)
