"""Oracle checks for beRTAMC2D's non-uniform temperatures: calculate_gradientT
(geometry_2d.cpp:524-576), the gradient sources (source_term_gradient,
beyondRTA_MC2d.cpp:597-676), the propagator chosen for a box temperature
(Scattering, beyondRTA_MC2d.cpp:1131-1166) and its interpolation (lirp,
collision_operator.cpp:918-929), and the host's sampling distribution of
propagator_H::get (bte_prop_pick_pmf) against the oracle's lower_bound."""
import numpy as np
import pytest

import oracle_lib as orc

KB = 1.3806488e-23


def grid_layout(nx, ny, dx=10.0, dy=7.0):
    rect = [[i * dx, (i + 1) * dx, j * dy, (j + 1) * dy] for j in range(ny) for i in range(nx)]
    return np.array(rect)


def centres(rect):
    return 0.5 * (rect[:, 0] + rect[:, 1]), 0.5 * (rect[:, 2] + rect[:, 3])


def test_gradient_of_a_linear_field_is_exact():
    rect = grid_layout(3, 3)
    cx, cy = centres(rect)
    teq = 300.0 + 0.013 * cx - 0.021 * cy
    g = orc.gradient_T(rect, np.zeros(9, np.int32), teq)
    assert np.allclose(g, [[0.013, -0.021]] * 9, atol=1e-12)


def test_gradient_of_a_ribbon_is_the_minimum_norm_solution():
    # neighbours only along x: g_y is undetermined and the minimum-norm
    # solution sets it to 0; reservoirs do not count as neighbours
    rect = np.array([[0, 10, 0, 5], [10, 20, 0, 5], [20, 30, 0, 5], [30, 40, 0, 5], [-10, 0, 0, 5],
                     [40, 50, 0, 5]], float)
    kind = np.array([0, 0, 0, 0, 1, 1], np.int32)
    teq = np.array([300.0, 301.0, 303.0, 306.0, 250.0, 400.0])
    g = orc.gradient_T(rect, kind, teq)
    # box 0: one regular neighbour (box 1); box 1: (301-300)/-10 and (303-301)/10
    # least squares over h = (-10, 10), df = (-1, 2): g = (10*2 + 10*1)/200
    assert np.allclose(g[:4, 1], 0.0, atol=1e-15)
    assert np.allclose(g[:4, 0], [0.1, 0.15, 0.25, 0.3], atol=1e-13)
    assert np.all(g[4:] == 0.0)


def test_gradient_matches_numpy_pseudo_inverse_on_irregular_layouts():
    # two columns of unequal boxes: T-junction edges and point contacts
    rect = np.array([[0, 10, 0, 40], [0, 10, 40, 100], [10, 25, 0, 30], [10, 25, 30, 70], [10, 25, 70, 100],
                     [25, 30, 0, 100], [-5, 0, 0, 100]], float)
    kind = np.array([0, 0, 0, 0, 0, 0, 1], np.int32)
    rng = np.random.default_rng(4)
    nc, nbr, _ = orc.contacts(rect)
    cx, cy = centres(rect)
    for trial in range(5):
        teq = 300.0 + rng.normal(0, 3, len(kind))
        g = orc.gradient_T(rect, kind, teq)
        for b in range(len(kind)):
            if kind[b] != 0:
                assert np.all(g[b] == 0.0)
                continue
            nb = [int(c) for c in nbr[b, :nc[b]] if kind[c] == 0]
            h = np.array([[cx[c] - cx[b], cy[c] - cy[b]] for c in nb])
            df = np.array([teq[c] - teq[b] for c in nb])
            ref = np.linalg.lstsq(h, df, rcond=None)[0]
            assert np.allclose(g[b], ref, rtol=1e-10, atol=1e-13), (trial, b, g[b], ref)


def _signed_cumm(P):
    """propagator_H's signed cumulative per column (collision_operator.cpp:857-870)"""
    a = np.abs(P)
    c = np.cumsum(a, axis=0) / a.sum(0)
    return np.where(P < 0, -c, c)


def test_propagator_pair_rules():
    rng = np.random.default_rng(2)
    n = 12
    Ps = [rng.normal(size=(n, n)) for _ in range(3)]
    Ts = [290.0, 300.0, 310.0]
    sc = [_signed_cumm(P) for P in Ps]
    # a tabulated temperature uses its own propagator
    for rule in (0, 1):
        assert np.allclose(orc.prop_interp(Ps, Ts, 300.0, rule), sc[1], atol=1e-15)
    # 295: Scattering takes the first key not below 295 (300) and the next
    # (310) and evaluates the line of their signed cumulatives at 295; the
    # bracketing rule interpolates the matrices at 290 and 300
    assert np.allclose(orc.prop_interp(Ps, Ts, 295.0, 0), sc[1] - 0.5 * (sc[2] - sc[1]), atol=1e-14)
    assert np.allclose(orc.prop_interp(Ps, Ts, 295.0, 1), _signed_cumm(0.5 * (Ps[0] + Ps[1])), atol=1e-14)
    # 305 lies between the last two keys: no key after 310 for rule 0
    with pytest.raises(ValueError):
        orc.prop_interp(Ps, Ts, 305.0, 0)
    assert np.allclose(orc.prop_interp(Ps, Ts, 302.5, 1), _signed_cumm(0.75 * Ps[1] + 0.25 * Ps[2]), atol=1e-14)
    for Tx in (280.0, 320.0):
        for rule in (0, 1):
            with pytest.raises(ValueError):
                orc.prop_interp(Ps, Ts, Tx, rule)
    # order of the input does not matter
    assert np.allclose(orc.prop_interp(Ps[::-1], Ts[::-1], 295.0, 1), _signed_cumm(0.5 * (Ps[0] + Ps[1])),
                       atol=1e-14)


def test_signed_cumulative_interpolation_is_unstable_under_sign_noise(pkg):
    """Why the bracketing rule interpolates matrices: Krylov propagators of
    the synthetic operator at 300 K and 306 K hold entries far below rounding
    level, whose computed signs are noise (modelled by adding 1e-17). Where
    an entry's sign differs between the two tables, lirp's cumulative dips
    and the next entry takes a large probability, negative when that entry
    is; the interpolated matrix keeps the negative weight at noise level."""
    mat = pkg.workloads.material(96, seed=2)
    rng = np.random.default_rng(0)
    Ps = []
    for T in (300.0, 306.0):
        B = orc.synthetic_B(KB * orc.bose_einstein_kernel(mat["omega"], T), mat["w"], 0.05, 2)
        Ps.append(orc.build_P(B, 5.0, 20, 0) + rng.normal(0, 1e-17, (96, 96)))
    a, b = _signed_cumm(Ps[0]), _signed_cumm(Ps[1])
    lirp = a + 0.5 * (b - a)
    mat_interp = orc.prop_interp(Ps, [300.0, 306.0], 303.0, 1)
    neg = lambda sc: max(pkg.prop_pick_pmf(sc[:, j])[sc[:, j] <= 0].sum() for j in range(96))
    assert neg(lirp) > 0.5        # supercritical sign-flip branching
    assert neg(mat_interp) < 1e-12


def test_get_on_a_signed_cumulative():
    rng = np.random.default_rng(5)
    R = (np.arange(20000) + 0.5) / 20000
    # monotone: the first index with |sc| >= R (numpy's searchsorted left)
    sc = _signed_cumm(rng.normal(size=(30, 1)))[:, 0]
    sg, idx = orc.propH_get_sc(sc, R)
    ref = np.minimum(np.searchsorted(np.abs(sc), R, side="left"), 29)
    assert np.array_equal(idx, ref)
    assert np.array_equal(sg, np.where(sc[idx] > 0, 1, -1))
    # non-monotone (extrapolated): the search result never decreases in R
    sc2 = 1.6 * _signed_cumm(rng.normal(size=(30, 1)))[:, 0] - 0.6 * _signed_cumm(rng.normal(size=(30, 1)))[:, 0]
    _, idx2 = orc.propH_get_sc(sc2, R)
    assert np.all(np.diff(idx2) >= 0)


def _grid_pmf(sc, M=400000):
    R = (np.arange(M) + 0.5) / M
    _, idx = orc.propH_get_sc(sc, R)
    return np.bincount(idx, minlength=len(sc)) / M


def test_host_pick_pmf_matches_the_oracle_search(pkg):
    rng = np.random.default_rng(7)
    cases = [
        np.array([0.5, 0.2, 0.6, 1.0]),
        np.array([0.0, 0.0, 1.5, 1.0, 1.0]),
        np.array([0.1, -0.3, 0.25, -0.9]),           # |last| < 1: the rest goes to the last index
        _signed_cumm(rng.normal(size=(40, 1)))[:, 0],
    ]
    for k in range(6):
        a = _signed_cumm(rng.normal(size=(40, 1)))[:, 0]
        b = _signed_cumm(rng.normal(size=(40, 1)))[:, 0]
        f = rng.uniform(-1.0, 1.0)
        cases.append(a + f * (b - a))
    for sc in cases:
        pmf = pkg.prop_pick_pmf(sc)
        ref = _grid_pmf(sc)
        assert abs(pmf.sum() - 1.0) < 1e-12
        assert np.abs(pmf - ref).max() < 1e-4, (sc, pmf, ref)


# ---- beRTAMC2D runs ----

def _material(pkg, n=120, seed=3):
    return pkg.workloads.material(n, seed=seed)


def test_gradient_sources_generate_the_nabla_T_rate(pkg):
    """P = identity and a box much larger than v dt: after k steps the box
    holds k steps of gradient particles, whose mean signed count and flux are
    dt vol kB/nq sum_i (-g.v_i) kernel_i (1, vx, vy) / Eeff, independent of
    Eeff (sampling.cpp:118-183)."""
    mat = _material(pkg)
    n = len(mat["omega"])
    L, thick, dt, T = 2.0e5, 1.0, 1.0, 300.0
    rect = np.array([[0.0, L, 0.0, L]])
    kind = np.array([0], np.int32)
    teq = np.array([T])
    g = (2.0e-3, -1.0e-3)
    ker = orc.bose_einstein_kernel(mat["omega"], T)
    vg = -(g[0] * mat["vx"] + g[1] * mat["vy"])
    vuc = mat["V"] * thick / mat["lattz"]
    vol = L * L * thick
    lam_E = dt * (vol / vuc) * KB * np.sum(np.abs(vg * ker))  # = Eeff * expected count
    Ed = lam_E / 1.0e5  # 1e5 particles per step
    res = orc.bert_run(mat, rect, kind, teq, teq, thickness=thick, dt=dt, maxtime=2.5 * dt, P=np.eye(n), Ed=Ed,
                       seed=11, gradient=g)
    assert res["nsteps"] == 3
    p = vg * ker  # signed pmf (unnormalised)
    pn = np.abs(p) / np.abs(p).sum()
    s = np.sign(p)
    scale = {0: 1e27 * Ed / vol, 1: 1e30 * Ed / vol, 2: 1e30 * Ed / vol}
    moments = {0: np.ones(n), 1: mat["vx"], 2: mat["vy"]}
    lam = 1.0e5
    for k in (1, 3):
        for q in (0, 1, 2):
            mean = k * lam * np.sum(pn * s * moments[q]) * scale[q]
            sd = np.sqrt(k * lam * np.sum(pn * moments[q] ** 2)) * scale[q]
            got = res["props"][k, 0, q]
            assert abs(got - mean) < 5 * sd, (k, q, got, mean, sd)
    # the flux follows -grad T: along +x for g_x < 0 would flip; here g_x > 0
    assert res["props"][3, 0, 1] < 0


def test_computed_gradients_equal_the_fixed_gradient_on_a_linear_ribbon(pkg):
    mat, B = pkg.workloads.config1_inputs(nmodes=150, seed=5)
    import scipy.linalg
    P = scipy.linalg.expm(2.0 * B)
    rect, kind, _, _ = pkg.workloads.ribbon(100.0, 40.0, 5, 300.0)
    cx = 0.5 * (rect[:, 0] + rect[:, 1])
    teq = 300.0 + 0.02 * cx
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 6, P=P, Efactor=50000, Tmax=302.0, Tmin=300.0, seed=3)
    a = orc.bert_run(pkg.workloads.material(150, seed=5), rect, kind, teq, teq, **kw)
    b = orc.bert_run(pkg.workloads.material(150, seed=5), rect, kind, teq, teq, gradient=(0.02, 0.0), **kw)
    assert np.array_equal(a["npart"], b["npart"])
    assert np.allclose(a["props"], b["props"], rtol=1e-12, atol=0)
    assert a["npart"][-1] > 1000
    # a uniform temperature has no gradient and no sources
    u = orc.bert_run(mat, rect, kind, np.full(5, 300.0), np.full(5, 300.0), **kw)
    assert np.all(u["npart"] == 0)


def _still_material(n=40, seed=8):
    """modes that do not move: scattering alone changes the histogram"""
    rng = np.random.default_rng(seed)
    return {"omega": rng.uniform(1.0, 10.0, n), "vx": np.zeros(n), "vy": np.zeros(n),
            "w": rng.uniform(0.1, 1.0, n), "V": 0.2, "lattz": 1.0}


def _shift(n, k):
    """P with column j = e_{min(j+k, n-1)}: nonnegative, and so are its
    extrapolations below used here"""
    P = np.zeros((n, n))
    for j in range(n):
        P[min(j + k, n - 1), j] = 1.0
    return P


def test_each_box_scatters_with_its_temperature_propagator():
    mat = _still_material()
    n = len(mat["omega"])
    rect = np.array([[0, 10, 0, 10], [10, 20, 0, 10]], float)
    kind = np.array([0, 0], np.int32)
    kw = dict(thickness=1.0, dt=1.0, maxtime=0.5, Efactor=20000, Tmax=312.0, Tmin=300.0, seed=21)
    Ts = [300.0, 310.0, 320.0]
    ident = [np.eye(n)] * 3
    shifts = [np.eye(n), _shift(n, 1), _shift(n, 2)]
    # tabulated temperatures: box 1 at 310 shifts every mode by one
    teq, tin = np.array([300.0, 310.0]), np.array([301.0, 311.0])
    a = orc.bert_run(mat, rect, kind, teq, tin, Ps=ident, Ts=Ts, **kw)
    b = orc.bert_run(mat, rect, kind, teq, tin, Ps=shifts, Ts=Ts, **kw)
    Ha, Hb = a["H_last"], b["H_last"]
    assert Ha[1].sum() > 1000 and Ha[0].sum() > 1000
    assert np.array_equal(Hb[0], Ha[0])
    assert np.array_equal(Hb[1], shifts[1] @ Ha[1])
    # 305 under Scattering's rule: the line through 310 and 320 at 305 picks
    # the 310 entry of every column (|sc| = 1.5 there, 1 after it)
    teq, tin = np.array([300.0, 305.0]), np.array([301.0, 306.0])
    a = orc.bert_run(mat, rect, kind, teq, tin, Ps=ident, Ts=Ts, **kw)
    b = orc.bert_run(mat, rect, kind, teq, tin, Ps=shifts, Ts=Ts, **kw)
    assert np.array_equal(b["H_last"][1], shifts[1] @ a["H_last"][1])
    # bracketing 300 and 310 at 305: each particle moves with probability 1/2
    c = orc.bert_run(mat, rect, kind, teq, tin, Ps=shifts, Ts=Ts, prop_rule=1, **kw)
    h = a["H_last"][1].astype(float)
    exp = 0.5 * (h + shifts[1] @ h)
    sd = np.sqrt(0.25 * (h + shifts[1] @ h))
    z = (c["H_last"][1] - exp) / np.where(sd > 0, sd, 1.0)
    assert np.abs(z).max() < 5.0
    assert c["H_last"][1].sum() == h.sum()
    # between the last two keys Scattering has no pair: an error
    with pytest.raises(RuntimeError, match="Extrapolation"):
        orc.bert_run(mat, rect, kind, np.array([300.0, 315.0]), np.array([301.0, 316.0]), Ps=shifts, Ts=Ts, **kw)
