"""Oracle checks for the RTAMC2D spectral decomposition (<spectral>,
RTA_MC2D.cpp:61-63, 718-766, 1930-1945, 1978-2052, 2349-2405): the Gamma
density against scipy, and the bookkeeping through a material whose modes
share one frequency, where every spectrum is the box's own response times
one Gamma profile."""
import numpy as np
import pytest
import scipy.stats

import oracle_lib as orc

HBAR = 1.054571726e-34


@pytest.mark.parametrize("k,theta,x", [(2.5, 1.3, 0.7), (40.0, 0.25, 9.0), (1.0e4, 1.0e-3, 10.05),
                                        (0.7, 2.0, 0.3), (250.0, 0.04, 9.5)])
def test_gamma_pdf_matches_scipy(k, theta, x):
    ref = scipy.stats.gamma.pdf(x, a=k, scale=theta)
    assert orc.gamma_pdf(k, theta, x) == pytest.approx(ref, rel=1e-9, abs=1e-300)


def one_frequency_material(n=64, omega=5.0, seed=4):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, n)
    v = rng.uniform(1.0, 6.0, n)
    return {"omega": np.full(n, omega), "vx": v * np.cos(ang), "vy": v * np.sin(ang),
            "w": rng.uniform(0.05, 0.5, n), "V": 0.15, "lattz": 1.0}


def ribbon4():
    rect = np.array([[0, 25, 0, 20], [25, 50, 0, 20], [50, 75, 0, 20], [75, 100, 0, 20], [-10, 0, 0, 20],
                     [100, 110, 0, 20]], float)
    kind = np.array([0, 0, 0, 0, 1, 1], np.int32)
    teq = np.array([300.0] * 4 + [300.5, 299.5])
    return rect, kind, teq


def test_single_frequency_spectra_are_the_response_times_one_profile():
    mat = one_frequency_material()
    rect, kind, teq = ribbon4()
    K, sigma, T0 = 40, 0.3, 300.0
    kw = dict(thickness=1.0, T0=T0, dt=10.0, maxtime=200.0, nparticles=20000, seed=5, sigma=sigma)
    r = orc.rta_run(mat, rect, kind, teq, spectral={"boxes": [1, 2], "nomega": K}, **kw)
    omega = 5.0
    grid = np.linspace(0.0, omega, K) + 0.5 * omega / K   # Eigen LinSpaced + delta/2
    w = scipy.stats.gamma.pdf(grid, a=omega ** 2 / sigma, scale=sigma / omega)
    cv = orc.calc_cv(mat["omega"], mat["V"], T0)
    for k, box in enumerate((1, 2)):
        sp = r["spectral"][k]
        assert np.allclose(sp["omega"], grid, rtol=1e-14)
        dT = r["T"][:, box] - T0
        assert np.abs(dT).max() > 1e-4
        assert np.allclose(sp["dT"], np.outer(dT, w), rtol=1e-8, atol=1e-12 * np.abs(dT).max() * w.max())
        jx = r["jx"][:, box]
        assert np.allclose(sp["jx"], np.outer(jx, w), rtol=1e-8, atol=1e-12 * np.abs(jx).max() * w.max())
        assert np.allclose(sp["jy"], np.outer(r["jy"][:, box], w), rtol=1e-8,
                           atol=1e-12 * np.abs(jx).max() * w.max())
        # fd_q: the energy density over hbar omega per q-point, in the same
        # units: (T - T0) Cv nq / (1e12 hbar omega)
        assert np.allclose(sp["fd"][:, 0], dT * cv / (1e12 * HBAR * omega), rtol=1e-8,
                           atol=1e-12 * np.abs(sp["fd"]).max())


def test_low_broadening_outliers_are_raised_to_the_floor():
    """sigma below exp(p25 - 1.5 (p75 - p25)) of log sigma is replaced by
    that floor (calc_percentiles_log, isotopic_scattering.cpp:31-50), others
    are kept: a run with an outlier equals the run with the floor value."""
    rng = np.random.default_rng(1)
    mat = one_frequency_material(n=40, seed=7)
    mat["omega"] = rng.uniform(2.0, 8.0, 40)
    sig = rng.uniform(0.2, 0.4, 40)
    ls = np.sort(np.log(sig))
    floor = np.exp(ls[25 * 40 // 100] - 1.5 * (ls[75 * 40 // 100] - ls[25 * 40 // 100]))
    out = sig.copy()
    out[3] = floor * 1e-3
    ls2 = np.sort(np.log(out))
    floor2 = np.exp(ls2[25 * 40 // 100] - 1.5 * (ls2[75 * 40 // 100] - ls2[25 * 40 // 100]))
    assert floor2 == floor  # one outlier does not move the quartiles here
    rect, kind, teq = ribbon4()
    kw = dict(thickness=1.0, T0=300.0, dt=10.0, maxtime=100.0, nparticles=5000, seed=3, sigma=0.3)
    a = orc.rta_run(mat, rect, kind, teq, spectral={"boxes": [0], "nomega": 20, "sigma": out}, **kw)
    fixed = out.copy()
    fixed[3] = floor
    b = orc.rta_run(mat, rect, kind, teq, spectral={"boxes": [0], "nomega": 20, "sigma": fixed}, **kw)
    # (threads sum the tallies in varying order: equal up to rounding)
    assert np.allclose(a["spectral"][0]["dT"], b["spectral"][0]["dT"], rtol=1e-10, atol=0)
    c = orc.rta_run(mat, rect, kind, teq, spectral={"boxes": [0], "nomega": 20, "sigma": sig}, **kw)
    assert not np.allclose(a["spectral"][0]["dT"], c["spectral"][0]["dT"], rtol=1e-6, atol=0)
