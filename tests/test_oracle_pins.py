"""Closed forms that pin oracle functions no other check fixes: the RTAMC2D
temperature magnitude, beRTAMC2D reservoir injection, the adiabatic-wall
re-emission, calc_Eeff, the scattering-rate factor, specular reflection and
the score-function derivative of the specularity.

Expectations are computed here in numpy from the reference's definitions
(wall emission pmf and flux, sampling.cpp:189-269; calc_Eeff,
beyondRTA_MC2d.cpp:100-128; post-processing, RTA_MC2D.cpp:2326-2435), not by
calling the oracle's own routines for the quantity under test. Monte Carlo
estimates are compared under the rule of stat_rules.py."""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as orc
from stat_rules import assert_3sigma, exact_z
from test_oracle_mc import bar, horizontal_material

KB = 1.3806488e-23
HBAR = 1.054571726e-34


def n_be(om, T):
    return 1.0 / np.expm1(1e12 * HBAR * om / (KB * T))


def be_kernel(om, T):
    x = 1e12 * HBAR * om / (2 * KB * T)
    return (x / np.sinh(x)) ** 2


def onesample_z(x, mu):
    """replicas x (R, ...) against a known mean: Student t with R - 1 degrees
    of freedom, as a normal-equivalent z"""
    x = np.asarray(x, float)
    R = x.shape[0]
    t = (x.mean(0) - mu) / (x.std(0, ddof=1) / np.sqrt(R))
    return np.sign(t) * stats.norm.isf(stats.t.sf(np.abs(t), R - 1))


def test_rta_ballistic_temperature_is_the_residence_time_closed_form():
    """Ballistic particles of mode m leave the hot left wall at t = 0 and
    cross the bar at v_x,m (v_y = 0: no wall contact). Up to time t a
    particle spends clamp(t - x_a/v_x, 0, dx/v_x) in the bin [x_a, x_a + dx],
    so the step response T(bin, t) - T0 = (J_L / dx) E_m[that time] /
    (Cv / vcorr), with modes drawn from the wall pmf v_x omega |n(T_L) -
    n(T0)| and J_L = 1e12 hbar sum(pmf) / (nq V_uc) (sampling.cpp:189-269,
    RTA_MC2D.cpp:2326-2435). The exact per-particle variance gives each of
    the 60 x 10 cells a z-score. The right reservoir is at T0 and emits
    nothing."""
    m = horizontal_material()
    dx, nb = 10.0, 10
    rect, kind, teq = bar(nbins=nb, dx=dx, TL=300.5, TR=300.0)
    N, dt, tmax = 40000, 5.0, 300.0
    out = orc.rta_run(m, rect, kind, teq, thickness=1.0, T0=300.0, dt=dt, maxtime=tmax, nparticles=N, seed=11,
                      ballistic=True, sigma=5.0)
    vx, om = m["vx"], m["omega"]
    vuc = m["V"] * 1.0 / m["lattz"]
    pm = np.where(vx > 0, vx * om * np.abs(n_be(om, 300.5) - n_be(om, 300.0)), 0.0)
    JL = 1e12 * HBAR * pm.sum() / vuc
    p = pm / pm.sum()
    cv = KB * be_kernel(om, 300.0).sum() / m["V"]
    vcorr = 1.0 / m["lattz"]
    nt = int(round(tmax / dt))
    t = dt * np.arange(1, nt + 1)[:, None, None]            # end of each cumulative row
    xa = dx * np.arange(nb)[None, :, None]
    with np.errstate(divide="ignore"):
        res = np.where(vx > 0, np.clip(t - xa / vx, 0.0, dx / vx), 0.0)   # (nt, nb, modes)
    EX = (res * p).sum(-1)
    VX = (res ** 2 * p).sum(-1) - EX ** 2
    scale = JL / dx / (cv / vcorr)
    z = exact_z(out["T"][:, :nb] - 300.0, scale * EX, scale * np.sqrt(VX / N))
    assert_3sigma(z, "ballistic T(x, t)")
    # the late rows are the steady state: every bin at J_L sum(p / v_x) / (Cv/vcorr)
    steady = JL * (p[vx > 0] / vx[vx > 0]).sum() / (cv / vcorr)
    assert out["T"][-1, :nb].mean() - 300.0 == pytest.approx(steady, rel=0.02)


def test_rta_scattering_rate_factor_scales_every_rate():
    """wscale multiplies every scattering rate: a run with wscale = 2 is the
    run of the material with twice the rates, the same random streams, so
    the outputs are identical on one thread (the BE_derivative pmf, C w,
    only changes scale; more threads would only reorder the sums)."""
    m = orc.synthetic_modes(300, 7)
    m["V"], m["lattz"] = 20.0, 1.0
    m2 = dict(m, w=m["w"] * 2.0)
    rect, kind, teq = bar(nbins=6, dx=20.0, H=30.0)
    kw = dict(thickness=1.0, T0=300.0, dt=10.0, maxtime=300.0, nparticles=3000, seed=4, sigma=2.0, nthreads=1)
    a = orc.rta_run(m, rect, kind, teq, wscale=2.0, **kw)
    b = orc.rta_run(m2, rect, kind, teq, **kw)
    assert a["events"] == b["events"] and a["crossings"] == b["crossings"]
    for k in ("T", "jx", "jy"):
        np.testing.assert_array_equal(a[k], b[k])
    c = orc.rta_run(m, rect, kind, teq, **kw)
    assert c["events"] < a["events"]


def _mirror_material(n=240, seed=3):
    m = orc.synthetic_modes(n, seed, mirror=True)
    m["V"], m["lattz"] = 20.0, 1.0
    return m


def test_rta_specular_channel_flux_is_exact():
    """Ballistic particles between specular (p = 1) top and bottom sides keep
    v_x and only flip v_y (mirror partner), so every particle crosses every
    bin once with displacement dx: the steady flux of every bin is the
    emitted power per cross-section (|J_L| + |J_R|) 1e30 exactly, as in the
    channel without walls. Diffuse sides (p = 0) send part of the particles
    back and carry less."""
    m = _mirror_material()
    rect, kind, teq = bar(nbins=8, dx=10.0, H=20.0)
    kw = dict(thickness=1.0, T0=300.0, dt=50.0, maxtime=3000.0, nparticles=20000, seed=5, ballistic=True,
              sigma=5.0)
    spec = orc.rta_run(m, rect, kind, teq, specularity=(0.0, 1.0, 1.0, 0.0), **kw)
    diff = orc.rta_run(m, rect, kind, teq, **kw)
    vuc = m["V"] / m["lattz"]
    JL = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 1.0, 0.0, vuc)[2]
    JR = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, -1.0, 0.0, vuc)[2]
    np.testing.assert_allclose(spec["jx"][-1, :8], (abs(JL) + abs(JR)) * 1e30, rtol=1e-9)
    assert diff["jx"][-1, :8].mean() < 0.9 * spec["jx"][-1, :8].mean()
    # partners exist for every mode of the table; the wrong axis has none
    with pytest.raises(RuntimeError, match="mirror"):
        m2 = dict(m, vx=m["vx"] + 1e-3 * np.arange(len(m["vx"])))   # break the v_x pairs only
        rect2 = np.array([[0.0, 20.0, 0.0, 20.0], [-5.0, 0.0, 0.0, 20.0], [20.0, 25.0, 0.0, 20.0]])
        # a closed top and bottom with a left void side: v_x partners are needed there
        orc.rta_run(m2, rect2[[0, 2]], np.array([0, 1], np.int32), np.array([300.0, 301.0]),
                    specularity=(1.0, 0.0, 0.0, 0.0), thickness=1.0, T0=300.0, dt=50.0, maxtime=500.0,
                    nparticles=2000, seed=1, ballistic=True, sigma=5.0)


def test_rta_score_derivative_matches_central_differences():
    """dT/dp from the score function (every particle carries the sum of
    d/dp log P(outcome) over its wall events: 1/p specular, -1/(1-p)
    diffuse) against central differences with common random numbers,
    Delta p = 0.05, on the mean temperature of the hot half and the flux of
    the bar. Eight seeds give paired differences whose mean must be within
    3 sigma of 0 (finite-difference bias O(Delta p^2) is far below the
    noise), while the derivative itself is many sigma away from 0."""
    m = _mirror_material(200, 9)
    rect, kind, teq = bar(nbins=8, dx=25.0, H=40.0)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=800.0, nparticles=20000, sigma=5.0)
    pt, pb, h = 0.3, 0.7, 0.05
    d_top, d_bot, g_top, g_bot = [], [], [], []
    for seed in range(8):
        g = orc.rta_run(m, rect, kind, teq, seed=seed, specularity=(0.0, pb, pt, 0.0), grad=True, **kw)
        f = lambda sb, st: orc.rta_run(m, rect, kind, teq, seed=seed, specularity=(0.0, sb, st, 0.0), **kw)
        Tp, Tm = f(pb, pt + h), f(pb, pt - h)
        Bp, Bm = f(pb + h, pt), f(pb - h, pt)
        hot = slice(0, 4)
        fd_t = (Tp["T"][-1, hot].mean() - Tm["T"][-1, hot].mean()) / (2 * h)
        fd_b = (Bp["T"][-1, hot].mean() - Bm["T"][-1, hot].mean()) / (2 * h)
        sc_t = g["dTdp"][2, -1, hot].mean()
        sc_b = g["dTdp"][1, -1, hot].mean()
        d_top.append(sc_t - fd_t)
        d_bot.append(sc_b - fd_b)
        g_top.append(sc_t)
        g_bot.append(sc_b)
    assert_3sigma(onesample_z(np.array([d_top, d_bot]).T, 0.0), "score - central difference")
    for gs in (g_top, g_bot):
        gs = np.array(gs)
        assert abs(gs.mean()) > 4 * gs.std(ddof=1) / np.sqrt(len(gs))
    # sides with p = 0 or 1 (or not differentiated) carry no derivative
    g = orc.rta_run(m, rect, kind, teq, seed=1, specularity=(0.0, 1.0, 0.3, 0.0), grad=True, **kw)
    assert np.abs(g["dTdp"][[0, 1, 3]]).max() == 0.0 and np.abs(g["dTdp"][2]).max() > 0


def test_bert_reservoir_injection_gives_the_ballistic_flux():
    """beRTAMC2D Isothermal_walls (beyondRTA_MC2d.cpp:446-585): each step
    injects Ntogenerate(dt L thickness |J| / Eeff) particles from each
    reservoir, advected U dt. With horizontal modes and P = I (no
    scattering) a particle of mode m stays dx / v_x in a bin, so the steady
    flux of every bin is (|J_L| + |J_R|) 1e30 W/m^2 in expectation (the
    injected power per cross-section), J from the wall pmf of
    sampling.cpp:189-269. 12 replicas, the flux averaged over steps 60-100
    of each, against the exact value."""
    m = horizontal_material(120, 8)
    n = len(m["omega"])
    rect, kind, teq = bar(nbins=10, dx=10.0, H=40.0)
    tin = teq.copy()
    vuc = m["V"] / m["lattz"]
    Jmode = lambda T, sgn: np.where(sgn * m["vx"] > 0,
                                    sgn * m["vx"] * m["omega"] * np.abs(n_be(m["omega"], T) - n_be(m["omega"], 300.0)),
                                    0.0)
    JL = 1e12 * HBAR * Jmode(300.5, 1).sum() / vuc
    JR = 1e12 * HBAR * Jmode(299.5, -1).sum() / vuc
    dt = 5.0
    Ed = dt * 40.0 * 1.0 * JL / 1500.0     # about 1500 particles per step from each wall
    P = np.eye(n)
    reps = []
    for seed in range(12):
        o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=dt, maxtime=500.0, P=P, Ed=Ed, seed=seed)
        reps.append(o["props"][60:, :10, 1].mean(0))
    z = onesample_z(np.array(reps), (JL + JR) * 1e30)
    assert_3sigma(z, "reservoir-driven flux")


def test_bert_adiabatic_box_keeps_the_outTref_distribution():
    """A closed box with adiabatic diffuse walls (beyondRTA_MC2d.cpp:281-435)
    initialised at T0 + 1 K with the outTref pmf omega |n(Tinit) - n(T0)|
    (sampling.cpp:271-330): at uniform temperature the flux absorbed by a
    wall gives back T_border = Tinit through the J(T) spline, and the wall
    re-emits with (n.v) omega |n(Tinit) - n(T0)|, which is detailed balance
    for that distribution when the mode set carries no net flux at
    equilibrium: every mode comes with its images under v_x -> -v_x and
    v_y -> -v_y, so at each wall the absorbed and emitted sums of
    (n.v) omega |n(Tinit) - n(T0)| agree. (A random-angle table does carry a
    net flux in that distribution, which is then not stationary.) After 30
    steps of ballistic flight (P = I) the live particles still follow the
    outTref pmf: fractions in 16 (speed, direction) classes over 8
    replicas, against the pmf."""
    b = orc.synthetic_modes(40, 21)
    m = {"omega": np.tile(b["omega"], 4), "w": np.tile(b["w"], 4),
         "vx": np.concatenate([b["vx"], -b["vx"], b["vx"], -b["vx"]]),
         "vy": np.concatenate([b["vy"], b["vy"], -b["vy"], -b["vy"]]), "V": 20.0, "lattz": 1.0}
    n = len(m["omega"])
    rect = np.array([[0.0, 100.0, 0.0, 100.0]])
    kind = np.array([0], np.int32)
    teq = np.array([300.0])
    tin = np.array([301.0])
    pmf = m["omega"] * np.abs(n_be(m["omega"], 301.0) - n_be(m["omega"], 300.0))
    pmf /= pmf.sum()
    sp = np.hypot(m["vx"], m["vy"])
    ang = np.arctan2(m["vy"], m["vx"])
    cls = np.digitize(sp, np.quantile(sp, [0.25, 0.5, 0.75])) * 4 + ((ang + np.pi) // (np.pi / 2)).astype(int) % 4
    expect = np.bincount(cls, weights=pmf, minlength=16)
    fr = []
    for seed in range(8):
        o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=5.0, maxtime=150.0, P=np.eye(n), Efactor=40000,
                         Tmax=301.0, Tmin=300.0, seed=seed)
        h = o["H_last"][0].astype(float)
        assert (h >= 0).all()
        fr.append(np.bincount(cls, weights=h, minlength=16) / h.sum())
    assert_3sigma(onesample_z(np.array(fr), expect), "mode classes after re-emission")


def test_calc_Eeff_closed_form():
    """calc_Eeff (beyondRTA_MC2d.cpp:100-128, selection 2026-2070): for every
    regular box vol/V_uc * sum_{omega != 0} 1e12 hbar omega (|n(Tmax) -
    n(Tref)| + |n(Tmin) - n(Tref)|) / nq / Efactor, Tref the mean Teq of the
    non-periodic boxes (reservoirs included), and the minimum over boxes."""
    m = orc.synthetic_modes(50, 2)
    m["V"], m["lattz"] = 7.0, 1.0
    m["omega"][3] = 0.0      # an omega = 0 mode is skipped
    rect = np.array([[0.0, 30.0, 0.0, 10.0], [30.0, 40.0, 0.0, 10.0], [-5.0, 0.0, 0.0, 10.0]])
    kind = np.array([0, 0, 1], np.int32)
    teq = np.array([300.0, 300.0, 306.0])
    tin = np.array([303.0, 300.0, 306.0])
    o = orc.bert_run(m, rect, kind, teq, tin, thickness=2.0, dt=5.0, maxtime=4.0, P=np.eye(50), Efactor=1000,
                     seed=1, max_steps=1)
    om = m["omega"][m["omega"] != 0]
    Tref = teq.mean()
    dE = (1e12 * HBAR * om * (np.abs(n_be(om, 306.0) - n_be(om, Tref)) + np.abs(n_be(om, 300.0) - n_be(om, Tref)))).sum()
    vuc = m["V"] * 2.0 / m["lattz"]
    vols = np.array([300.0, 100.0]) * 2.0 / vuc
    assert o["Eeff"] == pytest.approx((vols * dE / 1 / 1000).min(), rel=1e-12)
    # Tmax / Tmin given in the input replace the box extremes
    o2 = orc.bert_run(m, rect, kind, teq, tin, thickness=2.0, dt=5.0, maxtime=4.0, P=np.eye(50), Efactor=1000,
                      Tmax=310.0, Tmin=295.0, seed=1, max_steps=1)
    dE2 = (1e12 * HBAR * om * (np.abs(n_be(om, 310.0) - n_be(om, Tref)) + np.abs(n_be(om, 295.0) - n_be(om, Tref)))).sum()
    assert o2["Eeff"] == pytest.approx(vols.min() * dE2 / 1000, rel=1e-12)


def test_bert_given_initial_distribution():
    """<init_distribution> (beyondRTA_MC2d.cpp:142-193, ref_distribution
    sampling.cpp:70-114): a box given a signed distribution fd over the
    modes after the Gamma offset starts with Ntogenerate(Vbox/Vucell, Eeff) =
    floor or ceil of Vbox/Vucell sum|fd| / (Eeff nq) particles, mode j +
    ngamma with probability |fd_j| / sum|fd| and the sign of fd_j. Modes
    with v = 0 and P = I stay put, so the histogram after one step is the
    initial one. A box at Tinit != Teq adds its outTref particles on top:
    both sets are kept (the reference clears the given ones,
    beyondRTA_MC2d.cpp:2096-2105, 218)."""
    n, ng = 40, 3
    m = orc.synthetic_modes(n, 31)
    m["vx"] = np.zeros(n)
    m["vy"] = np.zeros(n)
    m["omega"][:ng] = 0.0   # the Gamma acoustic modes of the offset
    m["V"], m["lattz"] = 20.0, 1.0
    rng = np.random.default_rng(5)
    fd = rng.uniform(0.2, 1.0, n - ng) * np.where(rng.random(n - ng) < 0.4, -1.0, 1.0) * 1e-22
    rect = np.array([[0.0, 50.0, 0.0, 40.0], [50.0, 100.0, 0.0, 40.0]])
    kind = np.array([0, 0], np.int32)
    teq = np.array([300.0, 300.0])
    tin = np.array([300.0, 301.0])
    Ed = 2e-23
    P = np.eye(n - ng)
    vol = 50.0 * 40.0 * 1.0 / (m["V"] * 1.0 / m["lattz"])
    gen = vol * np.abs(fd).sum() / (Ed * 1)
    counts = []
    for seed in range(6):
        o = orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=1.0, maxtime=0.5, P=P, ngamma=ng, Ed=Ed,
                         seed=seed, max_steps=1, init_dist={0: fd})
        h = o["H_last"][0].astype(float)
        assert abs(h).sum() in (np.floor(gen), np.floor(gen) + 1)
        assert (h[:ng] == 0).all()
        nz = h[ng:] != 0
        assert (np.sign(h[ng:][nz]) == np.sign(fd[nz])).all()
        counts.append(np.abs(h[ng:]))
        assert o["H_last"][1].sum() > 0     # box 1's outTref particles are there too
    cnt = np.sum(counts, 0)
    p = np.abs(fd) / np.abs(fd).sum()
    assert stats.chisquare(cnt, cnt.sum() * p).pvalue > 1e-4
    with pytest.raises(RuntimeError, match="length"):
        orc.bert_run(m, rect, kind, teq, tin, thickness=1.0, dt=1.0, maxtime=0.5, P=P, ngamma=ng, Ed=Ed, seed=1,
                     max_steps=1, init_dist={0: fd[:-1]})
