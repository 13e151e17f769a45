"""Pulse spreading under RTA: the exact moments of beRTAMC2D's time-stepped
process (tests/msd_closed_form.py) against the reference's analytic1d MSD and
the textbook single-tau limit, and the oracle's beRTAMC2D (RTA scattering,
periodic ghost boxes) against those moments."""
import numpy as np
import pytest

import oracle_lib as orc
from msd_closed_form import (analytic1d_msd_laplace, be_kernel, bose_einstein, markov_msd, stehfest,
                             vacf_msd_laplace)


def msd_material(n=48, seed=31):
    """n/2 random modes and their inverses (-vx, -vy) with the same omega and
    tau, so that no spectrum carries a net velocity"""
    rng = np.random.default_rng(seed)
    h = n // 2
    om = 2 * np.pi * rng.uniform(1.0, 10.0, h)
    sp = rng.uniform(0.5, 4.0, h)
    ang = rng.uniform(0, 2 * np.pi, h)
    tau = 10 ** rng.uniform(np.log10(2.0), np.log10(50.0), h)
    vx, vy = sp * np.cos(ang), sp * np.sin(ang)
    om, tau = np.concatenate([om, om]), np.concatenate([tau, tau])
    C = be_kernel(om, 300.0)
    return {"omega": om, "vx": np.concatenate([vx, -vx]), "vy": np.concatenate([vy, -vy]), "w": 1.0 / tau,
            "V": 1.3806488e-23 * C.sum() / 1.6e-21, "lattz": 1.0}


def test_single_tau_limit_is_the_telegrapher_msd():
    """one relaxation time: MSD = 2 <v^2> tau (t - tau (1 - exp(-t/tau)))
    from both Laplace forms, and from the Markov moments as dt -> 0"""
    m = msd_material()
    tau = np.full(len(m["w"]), 7.0)
    C = be_kernel(m["omega"], 300.0)
    v2 = np.sum(C * m["vx"] ** 2) / C.sum()
    for t in (1.0, 5.0, 20.0, 60.0):
        exact = 2 * v2 * tau[0] * (t - tau[0] * (1 - np.exp(-t / tau[0])))
        for F in (analytic1d_msd_laplace(C, m["vx"], tau), vacf_msd_laplace(C, m["vx"], tau)):
            assert abs(stehfest(F, t) - exact) < 1e-5 * exact
    dt = 0.005
    mean, var = markov_msd(m["vx"], 1.0 / tau, C, C / tau, dt, int(60.0 / dt))
    for t in (5.0, 20.0, 60.0):
        exact = 2 * v2 * tau[0] * (t - tau[0] * (1 - np.exp(-t / tau[0])))
        k = int(round(t / dt))
        assert abs(var[k] + mean[k] ** 2 - exact) < 2e-3 * exact


def test_markov_moments_reach_the_reference_analytic1d_msd():
    """a pulse deposited with the re-emission spectrum C / tau: the moments of
    the time-stepped process converge to the reference's MSD(s) inverted by
    Gaver-Stehfest (error O(dt / tau)); with the equilibrium spectrum C they
    converge to the VACF form instead, which differs by more than 5 %"""
    m = msd_material()
    tau = 1.0 / m["w"]
    C = be_kernel(m["omega"], 300.0)
    dt = 0.01
    T = 60.0
    _, var_r = markov_msd(m["vx"], m["w"], C / tau, C / tau, dt, int(T / dt))
    _, var_c = markov_msd(m["vx"], m["w"], C, C / tau, dt, int(T / dt))
    F_ref = analytic1d_msd_laplace(C, m["vx"], tau)
    F_vacf = vacf_msd_laplace(C, m["vx"], tau)
    gap = 0.0
    for t in (2.0, 10.0, 30.0, 60.0):
        k = int(round(t / dt))
        ref, vac = stehfest(F_ref, t), stehfest(F_vacf, t)
        assert abs(var_r[k] - ref) < 5e-3 * ref, (t, var_r[k], ref)
        assert abs(var_c[k] - vac) < 5e-3 * vac, (t, var_c[k], vac)
        gap = max(gap, abs(ref - vac) / vac)
    assert gap > 0.05


def ribbon_row(nb=300, dx=2.0, H=20.0, pulse=150, Tinit=301.0):
    """nb boxes of dx x H along x, periodic in y through ghost boxes below and
    above the row (each translates onto the other), a pulse in box `pulse`."""
    x0 = -nb * dx / 2
    rect = [[x0 + k * dx, x0 + (k + 1) * dx, 0.0, H] for k in range(nb)]
    rect += [[x0, -x0, -2.0, 0.0], [x0, -x0, H, H + 2.0]]
    kind = np.array([0] * nb + [2, 2], np.int32)
    teq = np.full(nb + 2, 300.0)
    tin = teq.copy()
    tin[pulse] = Tinit
    periodic = {nb: (nb + 1, 0.0, H), nb + 1: (nb, 0.0, -H)}
    xc = np.array([0.5 * (r[0] + r[1]) for r in rect[:nb]])
    return np.array(rect), kind, teq, tin, periodic, xc


def moments_from_props(props, xc):
    """mean and grouped variance of the box energies at each step"""
    e = props[:, :len(xc), 0]
    tot = e.sum(1)
    mean = (e * xc).sum(1) / tot
    var = (e * (xc[None, :] - mean[:, None]) ** 2).sum(1) / tot
    return mean, var, tot


def expected_variance(m, dt, nsteps, dx, Tinit=301.0, T0=300.0):
    """variance of the box-grouped energy profile: displacement variance plus
    the uniform spread of the initial box and the grouping (Sheppard), dx^2/12
    each; the initial modes follow outTref, omega |n(Tinit) - n(T0)|
    (sampling.cpp:271-330), the redrawn ones BE_derivative at T0"""
    om = m["omega"]
    p0 = om * np.abs(bose_einstein(om, Tinit) - bose_einstein(om, T0))
    r = be_kernel(om, T0) * m["w"]
    _, var = markov_msd(m["vx"], m["w"], p0, r, dt, nsteps)
    return var + dx * dx / 6.0


def test_oracle_bert_rta_pulse_spreads_as_the_markov_moments():
    m = msd_material()
    rect, kind, teq, tin, per, xc = ribbon_row()
    dt, nsteps, N = 0.5, 60, 200000
    o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=dt, maxtime=dt * (nsteps - 1), rta=True,
                     Efactor=N, Tmax=301.0, Tmin=300.0, seed=5, max_steps=nsteps, periodic=per)
    mean, var, tot = moments_from_props(o["props"], xc)
    ev = expected_variance(m, dt, nsteps, 2.0)
    assert np.allclose(tot, tot[0], rtol=1e-12)  # no wall is reached, energy stays
    for n in (20, 40, 60):
        assert abs(var[n] - ev[n]) < 0.02 * ev[n], (n, var[n], ev[n])
