"""Whole-solver properties of the oracle's RTAMC2D and beRTAMC2D loops."""
import numpy as np
import pytest

import oracle_lib as orc


def horizontal_material(n=200, seed=5):
    """modes moving along x only (vy = 0): no contact with top/bottom walls"""
    m = orc.synthetic_modes(n, seed)
    sp = np.hypot(m["vx"], m["vy"])
    m["vx"] = np.where(m["vx"] >= 0, sp, -sp)
    m["vy"] = np.zeros(n)
    m["V"] = 20.0
    m["lattz"] = 1.0
    return m


def bar(nbins=10, dx=10.0, H=40.0, TL=300.5, TR=299.5):
    rect = [[k * dx, (k + 1) * dx, 0.0, H] for k in range(nbins)]
    rect += [[-dx, 0.0, 0.0, H], [nbins * dx, (nbins + 1) * dx, 0.0, H]]
    kind = [0] * nbins + [1, 1]
    teq = [300.0] * nbins + [TL, TR]
    return np.array(rect), np.array(kind, np.int32), np.array(teq)


def periodic_strip(width=20.0, nrows=5, dy=10.0, TB=300.5, TT=299.5):
    """One column of nrows boxes, periodic in x through ghost boxes on both
    sides (each translates onto the other, as the reference's beam.xml), with
    reservoirs below and above. Returns rect, kind, teq, periodic."""
    H = nrows * dy
    rect = [[0.0, width, k * dy, (k + 1) * dy] for k in range(nrows)]
    rect += [[-2.0, 0.0, 0.0, H], [width, width + 2.0, 0.0, H]]       # ghosts
    rect += [[-2.0, width + 2.0, -2.0, 0.0], [-2.0, width + 2.0, H, H + 2.0]]  # reservoirs
    kind = [0] * nrows + [2, 2, 1, 1]
    teq = [300.0] * (nrows + 2) + [TB, TT]
    left, right = nrows, nrows + 1
    periodic = {left: (right, width, 0.0), right: (left, -width, 0.0)}
    return np.array(rect), np.array(kind, np.int32), np.array(teq), periodic


def tilted_y_material(n=60, seed=13):
    """modes within 60 degrees of the y axis (|vy| >= |v|/2), |v| 2-8 nm/ps"""
    rng = np.random.default_rng(seed)
    om = 2 * np.pi * rng.uniform(0.5, 15.0, n)
    sp = rng.uniform(2.0, 8.0, n)
    ang = np.pi / 2 + rng.uniform(-np.pi / 3, np.pi / 3, n) + np.where(rng.random(n) < 0.5, 0.0, np.pi)
    return {"omega": om, "vx": sp * np.cos(ang), "vy": sp * np.sin(ang), "w": np.full(n, 0.01), "V": 20.0,
            "lattz": 1.0}


def test_rta_periodic_ballistic_flux_is_exact():
    """Ballistic particles from the bottom reservoir wrap around the periodic
    x direction until they reach the top: each crosses every row once with
    displacement dy, so the steady cumulative y flux of every row is
    (|J_B| + |J_T|) * 1e30 whatever the translations did to x."""
    m = tilted_y_material()
    rect, kind, teq, per = periodic_strip()
    out = orc.rta_run(m, rect, kind, teq, thickness=1.0, T0=300.0, dt=50.0, maxtime=1000.0, nparticles=20000,
                      seed=3, ballistic=True, sigma=5.0, periodic=per)
    vuc = m["V"] / m["lattz"]
    JB = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 0.0, 1.0, vuc)[2]
    JT = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, 0.0, -1.0, vuc)[2]
    np.testing.assert_allclose(out["jy"][-1, :5], (abs(JB) + abs(JT)) * 1e30, rtol=1e-9)
    assert np.abs(out["jy"][-1, 5:7]).max() == 0.0  # no time is spent in the ghost boxes


def test_bert_periodic_strip_conserves():
    """beRTAMC2D with periodic ghost boxes: the signed count (particles +
    pending wall counts) is conserved with adiabatic walls and P >= 0."""
    m, B, P = _bert_inputs(24, seed=6)
    rect, kind, teq, per = periodic_strip(TB=300.0, TT=300.0)
    kind = kind.copy()
    # drop the reservoirs: a closed strip, adiabatic walls below and above
    rect, kind, teq = rect[:7], kind[:7], teq[:7]
    tin = teq.copy()
    tin[0] = 310.0
    o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=2.0, maxtime=40.0, P=P, Efactor=3000, Tmax=310.0,
                     Tmin=300.0, seed=4, periodic=per)
    tot = o["npart"] + o["pending"]
    assert (tot == tot[0]).all() and tot[0] > 0


def test_rta_ballistic_flux_is_exact():
    """Ballistic horizontal particles cross every box once with displacement
    dx, so the steady cumulative flux of every box equals the total emitted
    wall power per unit cross-section: (|J_L| + |J_R|) * 1e30 W/m^2 (the
    particle energy Eeff = einfo/N of RTA_MC2D.cpp:2177)."""
    m = horizontal_material()
    rect, kind, teq = bar()
    out = orc.rta_run(m, rect, kind, teq, thickness=1.0, T0=300.0, dt=50.0, maxtime=1000.0, nparticles=20000,
                      seed=3, ballistic=True, sigma=5.0)
    vuc = m["V"] * 1.0 / m["lattz"]
    JL = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 1.0, 0.0, vuc)[2]
    JR = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, -1.0, 0.0, vuc)[2]
    np.testing.assert_allclose(out["jx"][-1, :10], (abs(JL) + abs(JR)) * 1e30, rtol=1e-9)
    assert np.abs(out["jy"]).max() == 0.0
    # each particle: one flight segment per box, absorbed once
    assert out["crossings"] == 20000 * 10


def test_rta_flux_conserved_across_bins_at_steady_state():
    m = orc.synthetic_modes(600, 3)
    C = orc.bose_einstein_kernel(m["omega"], 300.0)
    m["V"] = 1.3806488e-23 * C.sum() / 1.6e-21
    m["lattz"] = 1.0
    rect, kind, teq = bar(nbins=10, H=50.0)
    out = orc.rta_run(m, rect, kind, teq, thickness=1.0, T0=300.0, dt=100.0, maxtime=6000.0, nparticles=30000,
                      seed=9, sigma=1.0)
    jx = out["jx"][-1, :10]
    assert jx.std() / jx.mean() < 0.03
    T = out["T"][-1, :10]
    assert T[0] > 300.0 > T[-1]          # hot left, cold right
    assert abs((T[0] - 300) + (T[-1] - 300)) < 0.1 * (T[0] - T[-1])   # antisymmetric about T0


def _bert_inputs(n=40, seed=1, vzero=False):
    m = orc.synthetic_modes(n, seed)
    if vzero:
        m["vx"] = np.zeros(n)
        m["vy"] = np.zeros(n)
    C = orc.bose_einstein_kernel(m["omega"], 300.0)
    m["V"] = 1.3806488e-23 * C.sum() / 1.6e-21
    m["lattz"] = 1.0
    B = orc.synthetic_B(C, m["w"], 0.3, seed)
    P = orc.expm(5.0 * B)
    return m, B, P


def test_bert_conserves_particles_plus_wall_counts():
    m, B, P = _bert_inputs(60)
    rect = np.array([[k * 10.0, (k + 1) * 10.0, 0.0, 30.0] for k in range(5)])
    kind = np.zeros(5, np.int32)
    teq = np.full(5, 300.0)
    tin = teq.copy()
    tin[0] = 301.0
    o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=5.0, maxtime=95.0, P=P, Efactor=20000,
                     Tmax=301.0, Tmin=300.0, seed=2)
    tot = o["npart"] + o["pending"]
    assert (tot == tot[0]).all()
    assert 19000 < o["npart"][0] < 21000


def test_bert_mode_distribution_follows_P():
    """With v = 0 nothing moves, so E[H_k] = P^k H_0 for the (single) box."""
    m, B, P = _bert_inputs(30, seed=4, vzero=True)
    rect = np.array([[0.0, 10.0, 0.0, 10.0], [10.0, 20.0, 0.0, 10.0]])
    kind = np.zeros(2, np.int32)
    teq = np.full(2, 300.0)
    tin = np.array([310.0, 300.0])
    hs = []
    for seed in range(4):
        o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=5.0, maxtime=10.0, P=P, Efactor=200000,
                         Tmax=310.0, Tmin=300.0, seed=seed, max_steps=3)
        hs.append(o["H_last"][0].astype(float))
    h = np.mean(hs, 0)
    # initial distribution: outTref pmf ~ omega |n(310)-n(300)|
    n310 = orc.bose_einstein(m["omega"], 310.0)
    n300 = orc.bose_einstein(m["omega"], 300.0)
    h0 = m["omega"] * np.abs(n310 - n300)
    expect = np.linalg.matrix_power(P, 3) @ (h0 / h0.sum()) * h.sum()
    sig = np.sqrt(expect / 4 + 1)
    z = (h - expect) / sig
    assert np.abs(z).max() < 5, z


def test_bert_cancellation_keeps_net_counts():
    """A propagator with negative entries creates particles of both signs;
    after the first (cancelling) step the survivors per (box, mode) are
    exactly |h| (beyondRTA_MC2d.cpp:2342-2359)."""
    m, B, P = _bert_inputs(20, seed=6, vzero=True)
    Pn = P.copy()
    Pn[0, :] += Pn[1, :] + 0.1   # columns still sum to one ...
    Pn[1, :] = -0.1              # ... with a negative row 1
    rect = np.array([[0.0, 10.0, 0.0, 10.0], [10.0, 20.0, 0.0, 10.0]])
    kind = np.zeros(2, np.int32)
    teq = np.full(2, 300.0)
    tin = np.array([305.0, 300.0])
    o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=5.0, maxtime=10.0, P=Pn, Efactor=50000,
                     Tmax=305.0, Tmin=300.0, seed=1, max_steps=1)
    H = o["H_last"]
    assert (H < 0).any()
    assert o["npart"][1] == np.abs(H).sum()
    # the signed total is unbiased: E = sum over columns of (P_n h0) = N0
    assert abs(H.sum() / o["npart"][0] - 1) < 0.02


def periodic_cell(L=40.0, H=40.0, d=2.0):
    """one box periodic in x and y through four ghost boxes"""
    rect = np.array([[0.0, L, 0.0, H], [-d, 0.0, 0.0, H], [L, L + d, 0.0, H], [0.0, L, -d, 0.0], [0.0, L, H, H + d]])
    kind = np.array([0, 2, 2, 2, 2], np.int32)
    teq = np.full(5, 300.0)
    periodic = {1: (2, L, 0.0), 2: (1, -L, 0.0), 3: (4, 0.0, H), 4: (3, 0.0, -H)}
    return rect, kind, teq, periodic


def bulk_flux(m, g, T0=300.0):
    """kinetic-theory flux -kappa . g of the relaxation-time solution in an
    unbounded medium [W/m^2] for a gradient g [K/nm]"""
    C = 1.3806488e-23 * orc.bose_einstein_kernel(m["omega"], T0)
    tau = 1.0 / m["w"]
    kxx = (C * m["vx"] ** 2 * tau).sum() / m["V"]
    kxy = (C * m["vx"] * m["vy"] * tau).sum() / m["V"]
    kyy = (C * m["vy"] ** 2 * tau).sum() / m["V"]
    return -(kxx * g[0] + kxy * g[1]) * 1e30, -(kxy * g[0] + kyy * g[1]) * 1e30


def inversion_symmetric(m):
    """each mode paired with its v -> -v partner (then a gradient source
    emits as much + as - energy, so the reference's alternating signs of
    gradient particles are unbiased)"""
    out = {k: np.concatenate([m[k], m[k]]) for k in ("omega", "w")}
    out["vx"] = np.concatenate([m["vx"], -m["vx"]])
    out["vy"] = np.concatenate([m["vy"], -m["vy"]])
    out["V"] = 20.0
    out["lattz"] = 1.0
    return out


def test_rta_steady_periodic_cell_gives_bulk_flux():
    """With a homogeneous gradient and no walls (periodic in x and y) the
    steady branch must reproduce the bulk relaxation-time flux J = -kappa g
    (the first flights carry the whole mean flux; re-emitted particles carry
    none on average). Checked within 5 standard errors over 8 seeds, both
    components."""
    m = inversion_symmetric(orc.synthetic_modes(40, 21))
    rect, kind, teq, per = periodic_cell()
    g = (1.0e-3, 0.0)
    js = []
    for s in range(8):
        o = orc.rta_steady_run(m, rect, kind, teq, thickness=1.0, T0=300.0, gradient=g, nparticles=20000, seed=s,
                               sigma=5.0, periodic=per, max_generations=60)
        js.append((o["jx"][0], o["jy"][0]))
        assert o["generations"] >= 1
    js = np.array(js)
    for k, j0 in enumerate(bulk_flux(m, g)):
        se = js[:, k].std(ddof=1) / np.sqrt(len(js))
        assert abs(js[:, k].mean() - j0) < 5 * se, (k, js[:, k].mean(), j0, se)
        assert se < 0.05 * abs(j0)

