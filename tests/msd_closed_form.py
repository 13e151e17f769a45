"""Closed forms for the spread of an energy pulse under the relaxation-time
approximation (test infrastructure).

* markov_msd: exact mean and variance of the displacement of a particle of
  beRTAMC2D's time-stepped RTA process (beyondRTA_MC2d.cpp:983-1062 and
  708-970): each step advects by v_x dt, then the mode is redrawn with
  probability 1 - exp(-w dt) from the BE_derivative pmf kernel(omega, T) w
  (sampling.cpp:47-68). With p_k the mode distribution of step k and
  M[i, j] = (1 - q_i) delta_ij + q_i r_j, the partial moments
  E[X^a; mode = i] (a = 0, 1, 2) advance exactly: linear algebra, no sampling.
* analytic1d_msd_laplace: the reference's pulse MSD in Laplace space,
  MSD(s) = 2 sum C v^2 tau / (1 + s tau)^2 / (s^2 sum C / (1 + s tau))
  (analytic1d.cpp:1871-1913), and stehfest, the Gaver-Stehfest inversion
  with 16 terms that the reference applies to it (analytic1d.cpp:1918-1997),
  written from Stehfest's published coefficients.
* vacf_msd_laplace: the same for a pulse deposited with the equilibrium
  spectrum (modes ~ C), 2 sum C v^2 tau / (1 + s tau) / (s^2 sum C): the
  stationary velocity autocorrelation sum pi v^2 exp(-t / tau) integrated
  twice. The reference's form is that of a pulse deposited with the
  re-emission spectrum C / tau; both agree when every mode has one tau.
"""
from math import factorial, log

import numpy as np

HBAR = 1.05457172647e-22  # J ps
KB = 1.3806488e-23        # J / K


def bose_einstein(omega, T):
    return 1.0 / np.expm1(HBAR * omega / (KB * T))


def be_kernel(omega, T):
    x = HBAR * omega / (2 * KB * T)
    return (x / np.sinh(x)) ** 2


def markov_msd(vx, w, p0, r, dt, nsteps):
    """Mean and variance of the x displacement after n = 0..nsteps steps.
    Partial moments per mode (p, E[X; mode], E[X^2; mode]) are advected
    exactly, then mapped through M = diag(1 - q) + q r^T."""
    vx, w = np.asarray(vx, float), np.asarray(w, float)
    p = np.asarray(p0, float) / np.sum(p0)
    r = np.asarray(r, float) / np.sum(r)
    q = -np.expm1(-w * dt)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    mean = np.zeros(nsteps + 1)
    ex2 = np.zeros(nsteps + 1)
    step = lambda x: x * (1.0 - q) + (x @ q) * r
    for n in range(1, nsteps + 1):
        d = vx * dt
        m2 = m2 + 2.0 * d * m1 + d * d * p
        m1 = m1 + d * p
        mean[n], ex2[n] = m1.sum(), m2.sum()
        p, m1, m2 = step(p), step(m1), step(m2)
    return mean, ex2 - mean ** 2


def analytic1d_msd_laplace(C, vx, tau):
    C, vx, tau = (np.asarray(a, float) for a in (C, vx, tau))
    return lambda s: 2.0 * np.sum(C * vx ** 2 * tau / (1 + s * tau) ** 2) / (s * s * np.sum(C / (1 + s * tau)))


def vacf_msd_laplace(C, vx, tau):
    C, vx, tau = (np.asarray(a, float) for a in (C, vx, tau))
    return lambda s: 2.0 * np.sum(C * vx ** 2 * tau / (1 + s * tau)) / (s * s * np.sum(C))


def stehfest(F, t, N=16):
    """f(t) from its Laplace transform F by the Gaver-Stehfest formula."""
    h = N // 2
    V = []
    for k in range(1, N + 1):
        acc = 0.0
        for j in range((k + 1) // 2, min(k, h) + 1):
            acc += j ** h * factorial(2 * j) / (factorial(h - j) * factorial(j) * factorial(j - 1) *
                                              factorial(k - j) * factorial(2 * j - k))
        V.append((-1) ** (k + h) * acc)
    a = log(2.0) / t
    return a * sum(V[k - 1] * F(k * a) for k in range(1, N + 1))
