"""Size-independent properties of the HIP path, checked at the benchmark
sizes and on closed-form cases."""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib as orc
from stat_rules import assert_3sigma, exact_z
from test_oracle_mc import _bert_inputs, bar, horizontal_material

pytestmark = pytest.mark.gpu


def test_rta_ballistic_flux_is_exact(pkg, ctx):
    m = horizontal_material()
    rect, kind, teq = bar()
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    out = pkg.rtamc2d(ctx, modes, geom, thickness=1.0, T0=300.0, dt=50.0, maxtime=1000.0, nparticles=100000,
                      seed=5, ballistic=True, sigma_bounce=5.0)
    vuc = m["V"] / m["lattz"]
    JL = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 1.0, 0.0, vuc)[2]
    JR = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, -1.0, 0.0, vuc)[2]
    np.testing.assert_allclose(out["jx"][-1, :10], (abs(JL) + abs(JR)) * 1e30, rtol=1e-9)
    assert out["crossings"] == 100000 * 10
    assert np.abs(out["jy"]).max() == 0.0


@pytest.fixture(scope="module")
def config1(pkg, ctx):
    W = pkg.workloads
    mat, B = W.config1_inputs()
    P, rep = pkg.build_P(ctx, B, 5.0, m=30, precision="fp64", arnoldi="textbook")
    return mat, B, P, rep


def test_config1_propagator_matches_scipy(config1):
    mat, B, P, rep = config1
    E = scipy.linalg.expm(5.0 * B)
    assert np.abs(P - E).sum(0).max() < 1e-10
    assert rep["max_col_sum_err"] < 1e-12 and rep["breakdowns"] == 0


def test_config1_conserves_particles_at_full_size(pkg, ctx, config1):
    mat, B, P, _ = config1
    W = pkg.workloads
    c = W.config1()
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, c["rect"], c["kind"], c["Teq"], c["Tinit"])
    prop = pkg.Propagator(ctx, P)
    o = pkg.bertamc2d(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=c["maxtime"], max_steps=200,
                      Efactor=c["Efactor"], Tmax=301.0, Tmin=300.0, seed=77)
    assert o["nsteps"] == 200
    assert abs(o["npart"][0] / 1e6 - 1) < 0.005
    tot = o["npart"] + o["pending"]
    assert (tot == tot[0]).all()
    # energy density sums to the particle count (all particles positive)
    vol = (c["rect"][:, 1] - c["rect"][:, 0]) * (c["rect"][:, 3] - c["rect"][:, 2])
    count = (o["props"][:, :, 0] * vol[None, :] / (o["Eeff"] * 1e27)).sum(1)
    np.testing.assert_allclose(count, o["npart"], rtol=1e-9)
    # the pulse spreads: energy leaves bin 0 and reaches the far end
    ed = o["props"][:, :, 0]
    assert ed[-1, 0] < 0.2 * ed[0, 0] and ed[-1, -1] > 0


def test_bert_mode_distribution_follows_P(pkg, ctx):
    m, B, P = _bert_inputs(30, seed=4, vzero=True)
    rect = np.array([[0.0, 10.0, 0.0, 10.0], [10.0, 20.0, 0.0, 10.0]])
    kind = np.zeros(2, np.int32)
    teq = np.full(2, 300.0)
    tin = np.array([310.0, 300.0])
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    prop = pkg.Propagator(ctx, P)
    hs = []
    for seed in range(4):
        frames = []
        b = pkg.BertRun(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=10.0, max_steps=3, Efactor=200000,
                        Tmax=310.0, Tmin=300.0, seed=seed)
        b.execute(frame_cb=lambda s, t, e, H: frames.append(H))
        b.close()
        hs.append(frames[-1][0].astype(float))
    h = np.mean(hs, 0)
    n310 = orc.bose_einstein(m["omega"], 310.0)
    n300 = orc.bose_einstein(m["omega"], 300.0)
    h0 = m["omega"] * np.abs(n310 - n300)
    expect = np.linalg.matrix_power(P, 3) @ (h0 / h0.sum()) * h.sum()
    assert_3sigma(exact_z(h, expect, np.sqrt(expect / 4 + 1)), "P^3 H0 per mode")


def test_bert_cancellation_keeps_net_counts(pkg, ctx):
    m, B, P = _bert_inputs(20, seed=6, vzero=True)
    Pn = P.copy()
    Pn[0, :] += Pn[1, :] + 0.1
    Pn[1, :] = -0.1
    rect = np.array([[0.0, 10.0, 0.0, 10.0], [10.0, 20.0, 0.0, 10.0]])
    kind = np.zeros(2, np.int32)
    teq = np.full(2, 300.0)
    tin = np.array([305.0, 300.0])
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    prop = pkg.Propagator(ctx, Pn)
    frames = []
    b = pkg.BertRun(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=20.0, max_steps=3, Efactor=50000,
                    Tmax=305.0, Tmin=300.0, seed=1)
    o = b.execute(frame_cb=lambda s, t, e, H: frames.append(H))
    b.close()
    H0 = frames[0]
    assert (H0 < 0).any()
    # step 0 cancels: exactly |h| particles survive per (box, mode)
    assert o["npart"][1] == np.abs(H0).sum()
    assert abs(H0.sum() / o["npart"][0] - 1) < 0.02


def test_bert_capacity_overflow_is_an_error(pkg, ctx):
    m, B, P = _bert_inputs(20, seed=6, vzero=True)
    Pn = P.copy()
    Pn[0, :] += Pn[1, :] + 0.4
    Pn[1, :] = -0.4    # strong branching: the population grows every step
    rect = np.array([[0.0, 10.0, 0.0, 10.0], [10.0, 20.0, 0.0, 10.0]])
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, np.zeros(2, np.int32), np.full(2, 300.0), np.array([305.0, 300.0]))
    prop = pkg.Propagator(ctx, Pn)
    with pytest.raises(pkg.BteError, match="ECAPACITY"):
        pkg.bertamc2d(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=100.0, max_steps=20, Efactor=50000,
                      Tmax=305.0, Tmin=300.0, seed=1, pcancel_each=1000, capacity=1 << 21)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-5), ("bf16", 5e-5)])
def test_krylov_precisions_on_golden_operator(pkg, ctx, prec, tol):
    """BF16 within the tolerance that hi-only BF16 arithmetic misses by more
    than ten times (test_krylov_bf16_control.py)"""
    import os
    fix = np.load(os.path.join(os.path.dirname(__file__), "golden", "expm_fixtures.npz"))
    A, E = fix["ecop_n128_A"], fix["ecop_n128_E"]
    P, rep = pkg.build_P(ctx, A, 1.0, m=40, precision=prec, arnoldi="textbook")
    l1 = np.abs(P - E).sum(0).max()
    print(f"{prec}: max column L1 {l1:.3e}", flush=True)
    assert l1 < tol
    np.testing.assert_allclose(P.sum(0), 1.0, atol=1e-12)


def test_bert_replicas_in_one_geometry_are_isolated(pkg, ctx, config1):
    """Ribbon copies side by side in y share no side: a copy without an
    initial pulse stays empty while the pulsed copy evolves, so a run over R
    copies is R independent replicas."""
    mat, B, P, _ = config1
    W = pkg.workloads
    c = W.config1(replicas=2)
    tin = c["Tinit"].copy()
    tin[50:] = 300.0  # second copy: no pulse
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, c["rect"], c["kind"], c["Teq"], tin)
    nc, nbr, _ = geom.contacts()
    for b in range(100):  # no contact between the copies
        assert all((n >= 50) == (b >= 50) for n in nbr[b, :nc[b]])
    prop = pkg.Propagator(ctx, P)
    o = pkg.bertamc2d(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=c["maxtime"], max_steps=40,
                      Efactor=200000, Tmax=301.0, Tmin=300.0, seed=5)
    assert np.abs(o["props"][:, 50:, :]).max() == 0.0
    assert np.abs(o["props"][-1, :50, 0]).sum() > 0
    tot = o["npart"] + o["pending"]
    assert (tot == tot[0]).all()


@pytest.mark.parametrize("layout", ["ribbon", "zigzag", "many_modes"])
def test_bert_step_shapes_follow_the_same_histories(pkg, ctx, layout):
    """The step kernel with the mode velocities in LDS (one 1024-thread
    workgroup per CU) and with them in global memory (256-thread
    workgroups) differ only in which workgroup takes which particle. With a
    nonnegative propagator every particle keeps its sign, cancellation keeps
    every candidate, and the histories do not depend on that order: equal
    particle and pending counts and (box, mode) histograms at every step,
    box sums equal to rounding. 5000 modes do not fit the LDS next to the
    layout, so the automatic choice falls back to 256 threads."""
    from test_gpu_polygons import zigzag_ribbon
    W = pkg.workloads
    nmodes = 5000 if layout == "many_modes" else 400
    mat, B = W.config1_inputs(nmodes=nmodes, seed=3)
    P, _ = pkg.expm(ctx, B, 5.0)
    P = np.maximum(P, 0.0)
    P /= P.sum(0)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    if layout == "zigzag":
        polys = zigzag_ribbon(12, 120.0, 40.0, 3.0, 5.0)
        kind = np.zeros(len(polys), np.int32)
        teq = np.full(len(polys), 300.0)
        tin = teq.copy()
        tin[0] = 301.0
        geom = pkg.Geometry(ctx, None, kind, teq, tin, polys=polys)
    else:
        c = W.config1(nbins=20)
        geom = pkg.Geometry(ctx, c["rect"], c["kind"], c["Teq"], c["Tinit"])
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=5.0, maxtime=5.0 * 15, Efactor=100000, Tmax=301.0, Tmin=300.0, seed=9)
    out = []
    for shape in (0, 1):
        frames = []
        r = pkg.BertRun(ctx, modes, geom, prop, step_shape=shape, **kw)
        o = r.execute(frame_cb=lambda s, t, e, H: frames.append(H) and False)
        r.close()
        out.append((o, frames))
    (a, fa), (b, fb) = out
    print(f"{layout}: step threads {a['step_threads']} / {b['step_threads']}, "
          f"{a['particle_steps']} particle-steps, {len(fa)} frames", flush=True)
    assert a["step_threads"] == (256 if layout == "many_modes" else 1024)
    assert b["step_threads"] == 256
    assert a["nsteps"] == b["nsteps"] == 16
    np.testing.assert_array_equal(a["npart"], b["npart"])
    np.testing.assert_array_equal(a["pending"], b["pending"])
    assert a["particle_steps"] == b["particle_steps"]
    assert len(fa) == len(fb) > 0
    for x, y in zip(fa, fb):
        np.testing.assert_array_equal(x, y)
    scale = np.abs(a["props"]).max()
    assert scale > 0
    np.testing.assert_allclose(a["props"], b["props"], rtol=0, atol=1e-12 * scale)
