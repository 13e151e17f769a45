"""beRTAMC2D with non-uniform box temperatures on the GPU: gradient sources
(source_term_gradient, beyondRTA_MC2d.cpp:597-676, with calculate_gradientT,
geometry_2d.cpp:524-576) and the propagator of each box temperature
(Scattering, beyondRTA_MC2d.cpp:1131-1166; lirp, collision_operator.cpp:
918-929), against the oracle and closed forms."""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib as orc
from stat_rules import assert_3sigma, exact_z
from test_gpu_parity import zstats

pytestmark = pytest.mark.gpu

KB = 1.3806488e-23


def _modes(pkg, ctx, mat):
    return pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])


def test_geometry_gradients_match_the_oracle(pkg, ctx):
    rect = np.array([[0, 10, 0, 40], [0, 10, 40, 100], [10, 25, 0, 30], [10, 25, 30, 70], [10, 25, 70, 100],
                     [25, 30, 0, 100], [-5, 0, 0, 100]], float)
    kind = np.array([0, 0, 0, 0, 0, 0, 1], np.int32)
    teq = 300.0 + np.random.default_rng(3).normal(0, 3, len(kind))
    g = pkg.Geometry(ctx, rect, kind, teq).gradient_T()
    assert np.allclose(g, orc.gradient_T(rect, kind, teq), rtol=1e-12, atol=1e-15)


def test_fixed_gradient_sources_generate_the_nabla_T_rate(pkg, ctx):
    """P = identity in a box much larger than v dt: after k steps the box
    holds k steps of gradient particles with mean signed count and flux
    dt vol kB/nq sum_i (-g.v_i) kernel_i (1, vx, vy) / Eeff."""
    mat = pkg.workloads.material(120, seed=3)
    n = len(mat["omega"])
    L, thick, dt, T = 2.0e5, 1.0, 1.0, 300.0
    rect = np.array([[0.0, L, 0.0, L]])
    kind = np.array([0], np.int32)
    g = (2.0e-3, -1.0e-3)
    ker = orc.bose_einstein_kernel(mat["omega"], T)
    vg = -(g[0] * mat["vx"] + g[1] * mat["vy"])
    vuc = mat["V"] * thick / mat["lattz"]
    vol = L * L * thick
    lam = 1.0e5
    Ed = dt * (vol / vuc) * KB * np.sum(np.abs(vg * ker)) / lam
    geom = pkg.Geometry(ctx, rect, kind, [T])
    prop = pkg.Propagator(ctx, np.eye(n), T=T)
    res = pkg.bertamc2d(ctx, _modes(pkg, ctx, mat), geom, prop, thickness=thick, dt=dt, maxtime=2.5 * dt, Ed=Ed,
                        seed=5, gradient=g)
    assert res["nsteps"] == 3
    p = vg * ker
    pn, s = np.abs(p) / np.abs(p).sum(), np.sign(p)
    scale = {0: 1e27 * Ed / vol, 1: 1e30 * Ed / vol, 2: 1e30 * Ed / vol}
    moments = {0: np.ones(n), 1: mat["vx"], 2: mat["vy"]}
    zs = []
    for k in (1, 3):
        for q in (0, 1, 2):
            mean = k * lam * np.sum(pn * s * moments[q]) * scale[q]
            sd = np.sqrt(k * lam * np.sum(pn * moments[q] ** 2)) * scale[q]
            zs.append(exact_z(res["props"][k, 0, q], mean, sd))
    assert_3sigma(zs, "gradient-source moments")


def test_gradient_sources_match_the_oracle(pkg, ctx):
    """A ribbon with a linear Teq profile and adiabatic walls: the gradient
    sources of calculate_gradientT drive the run."""
    mat, B = pkg.workloads.config1_inputs(nmodes=200, seed=9)
    P = scipy.linalg.expm(2.0 * B)
    rect, kind, _, _ = pkg.workloads.ribbon(120.0, 40.0, 6, 300.0)
    cx = 0.5 * (rect[:, 0] + rect[:, 1])
    teq = 300.0 + 0.05 * cx
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, rect, kind, teq)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 15, Efactor=100000, Tmax=306.0, Tmin=300.0)
    R = 10
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=30 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, teq, P=P, seed=50 + r, **kw) for r in range(R)]
    assert g[0]["npart"][-1] > 5000
    zs = [zstats(np.array([x["props"][[4, -1], :, q] for x in g]), np.array([x["props"][[4, -1], :, q] for x in o]))
          for q in (0, 1)]
    for name, f in {"npart": lambda x: float(x["npart"][-1]),
                    "jx_mean": lambda x: x["props"][-1, :, 1].mean()}.items():
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    assert_3sigma(zs, "gradient sources vs oracle")
    # heat flows from the hot end (high x) to the cold end
    assert np.mean([x["props"][-1, :, 1].mean() for x in g]) < 0


def test_six_hundred_gradient_boxes_match_the_oracle(pkg, ctx):
    """A 600-box ribbon with a linear Teq profile: 600 gradient sources
    (more parts than source workgroups, which run contiguous ranges of them
    into a bounded number of buffer segments) against the oracle, every box
    at two times."""
    mat, B = pkg.workloads.config1_inputs(nmodes=200, seed=9)
    P = scipy.linalg.expm(2.0 * B)
    rect, kind, _, _ = pkg.workloads.ribbon(6000.0, 40.0, 600, 300.0)
    assert len(rect) == 600
    cx = 0.5 * (rect[:, 0] + rect[:, 1])
    teq = 300.0 + 0.002 * cx
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, rect, kind, teq)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 11, Efactor=100000, Tmax=312.0, Tmin=300.0)
    R = 8
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=70 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, teq, P=P, seed=90 + r, **kw) for r in range(R)]
    assert g[0]["npart"][-1] > 20000
    zs = [zstats(np.array([x["props"][[4, -1], :, q] for x in g]), np.array([x["props"][[4, -1], :, q] for x in o]))
          for q in (0, 1)]
    for name, f in {"npart": lambda x: float(x["npart"][-1]),
                    "jx_mean": lambda x: x["props"][-1, :, 1].mean()}.items():
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    print(assert_3sigma(zs, "600 gradient boxes vs oracle"))
    assert np.mean([x["props"][-1, :, 1].mean() for x in g]) < 0


def _still_material(n=40, seed=8):
    rng = np.random.default_rng(seed)
    return {"omega": rng.uniform(1.0, 10.0, n), "vx": np.zeros(n), "vy": np.zeros(n),
            "w": rng.uniform(0.1, 1.0, n), "V": 0.2, "lattz": 1.0}


def _shift(n, k):
    P = np.zeros((n, n))
    for j in range(n):
        P[min(j + k, n - 1), j] = 1.0
    return P


def _last_H(pkg, ctx, modes, geom, prop, **kw):
    frames = []
    b = pkg.BertRun(ctx, modes, geom, prop, **kw)
    try:
        b.execute(frame_cb=lambda s, t, e, H: frames.append(H) and 0)
    finally:
        b.close()
    return frames[-1]


def test_each_box_scatters_with_its_temperature_propagator(pkg, ctx):
    """Modes that do not move: one step of scattering maps the histogram by
    the box's propagator exactly (deterministic P) or in distribution."""
    mat = _still_material()
    n = len(mat["omega"])
    modes = _modes(pkg, ctx, mat)
    rect = np.array([[0, 10, 0, 10], [10, 20, 0, 10]], float)
    kind = np.array([0, 0], np.int32)
    kw = dict(thickness=1.0, dt=1.0, maxtime=0.5, Efactor=20000, Tmax=312.0, Tmin=300.0, seed=21)
    Ts = [300.0, 310.0, 320.0]
    ident = pkg.Propagator(ctx, [np.eye(n)] * 3, T=Ts)
    shifts = [np.eye(n), _shift(n, 1), _shift(n, 2)]
    sprop = pkg.Propagator(ctx, shifts, T=Ts)
    g1 = pkg.Geometry(ctx, rect, kind, [300.0, 310.0], [301.0, 311.0])
    Ha = _last_H(pkg, ctx, modes, g1, ident, **kw)
    Hb = _last_H(pkg, ctx, modes, g1, sprop, **kw)
    assert Ha[0].sum() > 1000 and Ha[1].sum() > 1000
    assert np.array_equal(Hb[0], Ha[0])
    assert np.array_equal(Hb[1], shifts[1] @ Ha[1])
    # 305, Scattering's rule: extrapolating 310 and 320 picks the 310 entry
    g2 = pkg.Geometry(ctx, rect, kind, [300.0, 305.0], [301.0, 306.0])
    Ha = _last_H(pkg, ctx, modes, g2, ident, **kw)
    Hb = _last_H(pkg, ctx, modes, g2, sprop, **kw)
    assert np.array_equal(Hb[1], shifts[1] @ Ha[1])
    # bracketing 300 and 310: each particle moves with probability 1/2
    Hc = _last_H(pkg, ctx, modes, g2, sprop, prop_interp=1, **kw)
    h = Ha[1].astype(float)
    exp = 0.5 * (h + shifts[1] @ h)
    sd = np.sqrt(0.25 * (h + shifts[1] @ h))
    assert_3sigma(exact_z(Hc[1], exp, sd), "bracketing interpolation, binomial per mode")
    assert Hc[1].sum() == h.sum()
    # between the last two tabulated temperatures Scattering has no pair
    g3 = pkg.Geometry(ctx, rect, kind, [300.0, 315.0], [301.0, 316.0])
    with pytest.raises(pkg.BteError, match="Extrapolation"):
        pkg.BertRun(ctx, modes, g3, sprop, **kw)
    # a single propagator tabulated at 300 does not cover a box at 310
    p300 = pkg.Propagator(ctx, np.eye(n), T=300.0)
    with pytest.raises(pkg.BteError, match="Extrapolation"):
        pkg.BertRun(ctx, modes, g1, p300, **kw)


def test_interpolated_propagators_match_the_oracle(pkg, ctx):
    """Propagators at four temperatures, box temperatures between them, with
    gradient sources: Scattering's pair rule and the bracketing rule."""
    mat, B = pkg.workloads.config1_inputs(nmodes=160, seed=4)
    Ts = [300.0, 304.0, 308.0, 312.0]
    Ps = [scipy.linalg.expm((1.5 + 0.4 * k) * B) for k in range(4)]
    rect, kind, _, _ = pkg.workloads.ribbon(80.0, 40.0, 4, 300.0)
    teq = np.array([300.0, 302.0, 305.0, 307.5])
    tin = teq + np.array([1.0, 0.0, 0.0, 0.0])
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    prop = pkg.Propagator(ctx, Ps, T=Ts)
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 12, Efactor=60000, Tmax=308.0, Tmin=300.0)
    R = 10
    for rule in (0, 1):
        g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=70 + r, prop_interp=rule, **kw) for r in range(R)]
        o = [orc.bert_run(mat, rect, kind, teq, tin, Ps=Ps, Ts=Ts, prop_rule=rule, seed=90 + r, **kw)
             for r in range(R)]
        zs = [zstats(np.array([x["props"][[3, -1], :, q] for x in g]),
                     np.array([x["props"][[3, -1], :, q] for x in o])) for q in (0, 1)]
        zs.append(zstats(np.array([float(x["npart"][-1]) for x in g]), np.array([float(x["npart"][-1]) for x in o])))
        assert_3sigma(zs, f"interpolated propagators, rule {rule}")
