"""Batched RTAMC2D runs in one launch, graph replay, the launch-shape tuner,
specular edges and the specularity gradient (score function against
central finite differences with common random numbers)."""
import numpy as np
import pytest

import oracle_lib as orc
from stat_rules import assert_3sigma
from test_oracle_pins import onesample_z
from test_oracle_mc import bar

pytestmark = pytest.mark.gpu


def tilted_mirror_material(n=40, seed=11):
    """Mirror pairs (vx, vy), (vx, -vy) at angles within 60 degrees of the x
    axis, so |vx| >= |v|/2 and every ballistic particle crosses a short bar
    quickly; scattering rates are irrelevant in ballistic runs."""
    rng = np.random.default_rng(seed)
    h = n // 2
    om = 2 * np.pi * rng.uniform(0.5, 15.0, h)
    sp = rng.uniform(2.0, 8.0, h)
    ang = rng.uniform(-np.pi / 3, np.pi / 3, h) + np.where(rng.random(h) < 0.5, 0.0, np.pi)
    vx, vy = sp * np.cos(ang), sp * np.sin(ang)
    m = {"omega": np.concatenate([om, om]), "vx": np.concatenate([vx, vx]), "vy": np.concatenate([vy, -vy]),
         "w": np.full(n, 0.01), "V": 20.0, "lattz": 1.0}
    return m


def ribbon_runs(pkg, ctx, W, specs):
    out = []
    for (length, width, T, N, seed) in specs:
        rect, kind, teq, _ = W.ribbon(length, width, 8, T, reservoirs=(T + 0.5, T - 0.5))
        out.append(dict(geom=pkg.Geometry(ctx, rect, kind, teq), T0=T, nparticles=N, seed=seed))
    return out


def test_batch_equals_single_runs(pkg, ctx):
    W = pkg.workloads
    mat = W.material(300, seed=3)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    runs = ribbon_runs(pkg, ctx, W, [(80.0, 20.0, 300.0, 20000, 1), (60.0, 40.0, 250.0, 5000, 2),
                                      (100.0, 10.0, 400.0, 12000, 3)])
    kw = dict(thickness=1.0, dt=20.0, maxtime=200.0, sigma_bounce=2.0)
    b = pkg.RtaBatch(ctx, modes, runs, **kw)
    res, rep = b.execute()
    b.close()
    for r, o in zip(runs, res):
        s = pkg.rtamc2d(ctx, modes, r["geom"], T0=r["T0"], nparticles=r["nparticles"], seed=r["seed"], **kw)
        np.testing.assert_allclose(o["T"], s["T"], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(o["jx"], s["jx"], rtol=1e-9, atol=1e-9 * np.abs(s["jx"]).max())
    assert rep["events"] > 0


def test_graph_replay_and_tuner_keep_results(pkg, ctx):
    W = pkg.workloads
    mat = W.material(300, seed=3)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    runs = ribbon_runs(pkg, ctx, W, [(80.0, 20.0, 300.0, 20000, 1), (60.0, 40.0, 250.0, 5000, 2)])
    kw = dict(thickness=1.0, dt=20.0, maxtime=200.0, sigma_bounce=2.0)
    eager = pkg.RtaBatch(ctx, modes, runs, **kw)
    ref, _ = eager.execute()
    graph = pkg.RtaBatch(ctx, modes, runs, use_graph=True, **kw)
    for rep_i in range(3):
        got, _ = graph.execute()
        for a, b in zip(got, ref):
            np.testing.assert_allclose(a["T"], b["T"], rtol=1e-12, atol=1e-12)
        eager.execute()  # eager work between replays
        ctx.sync()
    blk, nb = graph.tune(fraction=0.2)
    assert blk in (128, 256, 512) and nb > 0
    got, rep = graph.execute()
    assert (rep["block"], rep["blocks"]) == (blk, nb)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a["T"], b["T"], rtol=1e-12, atol=1e-12)
    eager.close()
    graph.close()


def test_specular_channel_flux_is_exact(pkg, ctx):
    """Ballistic particles between specular (p = 1) walls keep vx, so each
    crosses every box once with displacement dx: the steady cumulative flux
    is (|J_L| + |J_R|) * 1e30 in every box, as for a channel without walls.
    Diffuse walls (p = 0) send part of it back."""
    m = tilted_mirror_material()
    rect, kind, teq = bar()
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    vuc = m["V"] / m["lattz"]
    JL = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 1.0, 0.0, vuc)[2]
    JR = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, -1.0, 0.0, vuc)[2]
    base = dict(geom=geom, T0=300.0, nparticles=50000, seed=4, ballistic=True)
    b = pkg.RtaBatch(ctx, modes, [dict(base, specularity=(0, 1, 1, 0)), dict(base, specularity=(0, 0, 0, 0))],
                     thickness=1.0, dt=50.0, maxtime=1000.0, sigma_bounce=5.0)
    (spec, diff), _ = b.execute()
    b.close()
    np.testing.assert_allclose(spec["jx"][-1, :10], (abs(JL) + abs(JR)) * 1e30, rtol=1e-9)
    assert diff["jx"][-1, :10].mean() < 0.9 * spec["jx"][-1, :10].mean()


def test_specularity_gradient_matches_finite_differences(pkg, ctx):
    """dT/dp_top and dT/dp_bottom from the score-function tallies against
    central differences (dp = 0.05) with common random numbers, over R seeds:
    the paired differences (score - central difference, same seeds) of the
    mean bin temperature derivative pass the 3-sigma rule on both sides, and
    the derivatives are nonzero."""
    m = tilted_mirror_material(n=80, seed=12)
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(200.0, 20.0, 8, 300.0, reservoirs=(301.0, 299.0))
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    p = np.array([0.0, 0.7, 0.3, 0.0])
    dp = 0.05
    R, N = 8, 200000
    runs = []
    for s in range(R):
        base = dict(geom=geom, T0=300.0, nparticles=N, seed=100 + s, ballistic=True)
        runs.append(dict(base, specularity=tuple(p), grad=True))
        for side in (1, 2):
            for sgn in (+1, -1):
                q = p.copy()
                q[side] += sgn * dp
                runs.append(dict(base, specularity=tuple(q)))
    b = pkg.RtaBatch(ctx, modes, runs, thickness=1.0, dt=20.0, maxtime=400.0, sigma_bounce=5.0)
    res, _ = b.execute(grad=True)
    b.close()
    # observable: mean temperature of the left half at the last time
    obs = lambda T: T[-1, :4].mean()
    zs = []
    for si, side in enumerate((1, 2)):
        sf, fd = [], []
        for s in range(R):
            g = res[5 * s]
            sf.append(obs(g["dTdp"][side]))
            plus, minus = res[5 * s + 1 + 2 * si], res[5 * s + 2 + 2 * si]
            fd.append((obs(plus["T"]) - obs(minus["T"])) / (2 * dp))
        sf, fd = np.array(sf), np.array(fd)
        se = np.sqrt(sf.var(ddof=1) / R + fd.var(ddof=1) / R)
        print(f"side {side}: score {sf.mean():.4e} +- {sf.std(ddof=1)/np.sqrt(R):.1e}, "
              f"FD {fd.mean():.4e} +- {fd.std(ddof=1)/np.sqrt(R):.1e}", flush=True)
        zs.append(onesample_z((sf - fd)[:, None], 0.0))
        assert abs(sf.mean()) > 4 * sf.std(ddof=1) / np.sqrt(R)
    assert_3sigma(zs, "score - central difference")


def test_fp32_lds_tallies_match_fp64(pkg, ctx):
    """Tally mode 1 (FP32 partial sums per workgroup in LDS, FP64 final sum)
    changes only the rounding: the particles draw the same streams, so the
    event counts are equal and T, jx, jy and dT/dp agree with the FP64
    device-atomic tallies to 1e-4 of each output's largest magnitude, also
    on graph replays with eager work between them."""
    m = tilted_mirror_material(n=80, seed=12)
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(200.0, 20.0, 8, 300.0, reservoirs=(301.0, 299.0))
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    runs = [dict(geom=geom, T0=300.0, nparticles=100000, seed=7, specularity=(0.0, 0.7, 0.3, 0.0), grad=True),
            dict(geom=geom, T0=300.0, nparticles=60000, seed=8, specularity=(0.0, 0.5, 0.5, 0.0))]
    kw = dict(thickness=1.0, dt=20.0, maxtime=400.0, sigma_bounce=5.0)
    ref_b = pkg.RtaBatch(ctx, modes, runs, **kw)
    ref, rrep = ref_b.execute(grad=True)
    lds = pkg.RtaBatch(ctx, modes, runs, use_graph=True, tally="fp32-lds", **kw)
    for replay in range(2):
        got, grep = lds.execute(grad=True)
        assert (grep["events"], grep["crossings"]) == (rrep["events"], rrep["crossings"])
        assert grep["block"] == 768 and grep["blocks"] == rrep["blocks"] // 8
        for a, b in zip(got, ref):
            for key, off in (("T", 300.0), ("jx", 0.0), ("jy", 0.0)):
                scale = np.abs(b[key] - off).max()
                assert scale > 0
                np.testing.assert_allclose(a[key] - off, b[key] - off, rtol=0, atol=1e-4 * scale)
            scale = np.abs(b["dTdp"]).max()
            np.testing.assert_allclose(a["dTdp"], b["dTdp"], rtol=0, atol=1e-4 * max(scale, 1e-30))
        assert np.abs(got[0]["dTdp"][1]).max() > 0 and np.abs(got[1]["dTdp"]).max() == 0
        ref_b.execute(grad=True)  # eager work between replays
        ctx.sync()
    ref_b.close()
    lds.close()


def test_fp32_lds_tallies_reject_what_does_not_fit(pkg, ctx):
    """600 regular boxes x 100 time bins x 3 arrays of FP32 exceed the LDS of
    a workgroup: the mode is refused before any launch and the batch keeps
    working with FP64 tallies."""
    W = pkg.workloads
    mat = W.material(64, seed=3)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    rect, kind, teq, _ = W.ribbon(6000.0, 20.0, 600, 300.0, reservoirs=(300.5, 299.5))
    geom = pkg.Geometry(ctx, rect, kind, teq)
    runs = [dict(geom=geom, T0=300.0, nparticles=2000, seed=1)]
    kw = dict(thickness=1.0, dt=10.0, maxtime=1000.0, sigma_bounce=2.0)
    with pytest.raises(Exception, match="LDS"):
        pkg.RtaBatch(ctx, modes, runs, tally="fp32-lds", **kw)
    b = pkg.RtaBatch(ctx, modes, runs, **kw)
    res, rep = b.execute()
    b.close()
    assert rep["events"] > 0 and np.isfinite(res[0]["T"]).all()



def test_fp64_lds_tallies_follow_the_same_histories(pkg, ctx):
    """Tally mode 2 (FP64 partial sums of T, jx and jy per workgroup in LDS for
    the leading time bins, device atomics for the rest: here 85 of 200 bins
    of 80 boxes fit) runs the same histories as mode 0: equal event and
    crossing counts, and T, jx, jy equal up to the order of the FP64 sums,
    also on graph replays; it refuses dT/dp before any launch."""
    W = pkg.workloads
    mat = W.material(300, seed=3, mirror=True)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    rect, kind, teq, _ = W.ribbon(400.0, 40.0, 40, 300.0, reservoirs=(300.5, 299.5))
    geom = pkg.Geometry(ctx, rect, kind, teq)
    runs = [dict(geom=geom, T0=300.0, nparticles=60000, seed=21),
            dict(geom=geom, T0=300.0, nparticles=30000, seed=22, specularity=(0.0, 1.0, 1.0, 0.0))]
    kw = dict(thickness=1.0, dt=5.0, maxtime=1000.0, sigma_bounce=2.0)
    ref_b = pkg.RtaBatch(ctx, modes, runs, **kw)
    ref, rrep = ref_b.execute()
    lds = pkg.RtaBatch(ctx, modes, runs, use_graph=True, tally="fp64-lds", **kw)
    for replay in range(2):
        got, grep = lds.execute()
        assert (grep["events"], grep["crossings"]) == (rrep["events"], rrep["crossings"])
        assert grep["block"] == 768
        for a, b in zip(got, ref):
            for key, off in (("T", 300.0), ("jx", 0.0), ("jy", 0.0)):
                scale = np.abs(b[key] - off).max()
                assert scale > 0
                np.testing.assert_allclose(a[key] - off, b[key] - off, rtol=0, atol=1e-10 * scale)
        ref_b.execute()  # eager work between replays
        ctx.sync()
    ref_b.close()
    lds.close()
    with pytest.raises(Exception, match="dT/dp"):
        pkg.RtaBatch(ctx, modes, [dict(runs[0], specularity=(0.0, 0.5, 0.5, 0.0), grad=True)],
                     tally="fp64-lds", **kw)


def test_fp64_run_lds_tallies_follow_the_same_histories(pkg, ctx):
    """Tally mode 3 (FP64 LDS sums per work item of one run's particles, the
    block holding that run's boxes only) on a batch of five runs of different
    geometries (10 to 80 bins), temperatures, rate factors, specular sides
    and particle counts (one run below one item of 4096 ids, others split in
    up to 30 items): the same histories as mode 0 (equal event and crossing
    counts, T, jx, jy equal up to the order of the FP64 sums) on two graph
    replays; dT/dp is refused before any launch."""
    W = pkg.workloads
    mat = W.material(300, seed=3, mirror=True)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    specs = [(400.0, 40.0, 40, 300.0, 60000, 21, 1.0, (0.0, 0.0, 0.0, 0.0)),
             (200.0, 20.0, 20, 150.0, 5000, 22, 0.5, (0.0, 0.0, 0.0, 0.0)),
             (800.0, 100.0, 80, 450.0, 120000, 23, 1.5, (0.0, 0.0, 0.0, 0.0)),
             (100.0, 50.0, 10, 300.0, 700, 24, 1.0, (0.0, 0.0, 0.0, 0.0)),
             (400.0, 40.0, 40, 300.0, 30000, 25, 1.0, (0.0, 1.0, 1.0, 0.0))]
    runs = []
    for (length, width, nb, T, N, seed, ws, sp) in specs:
        rect, kind, teq, _ = W.ribbon(length, width, nb, T, reservoirs=(T + 0.5, T - 0.5))
        runs.append(dict(geom=pkg.Geometry(ctx, rect, kind, teq), T0=T, nparticles=N, seed=seed, wscale=ws,
                         specularity=sp))
    kw = dict(thickness=1.0, dt=5.0, maxtime=1000.0, sigma_bounce=2.0)
    ref_b = pkg.RtaBatch(ctx, modes, runs, **kw)
    ref, rrep = ref_b.execute()
    lds = pkg.RtaBatch(ctx, modes, runs, use_graph=True, tally="fp64-lds-runs", **kw)
    for replay in range(2):
        got, grep = lds.execute()
        print(f"replay {replay}: events {grep['events']} crossings {grep['crossings']}", flush=True)
        assert (grep["events"], grep["crossings"]) == (rrep["events"], rrep["crossings"])
        assert grep["block"] == 768
        for r, (a, b) in enumerate(zip(got, ref)):
            T0 = specs[r][3]
            for key, off in (("T", T0), ("jx", 0.0), ("jy", 0.0)):
                scale = np.abs(b[key] - off).max()
                assert scale > 0, (r, key)
                np.testing.assert_allclose(a[key] - off, b[key] - off, rtol=0, atol=1e-10 * scale)
        ref_b.execute()  # eager work between replays
        ctx.sync()
    ref_b.close()
    lds.close()
    with pytest.raises(Exception, match="dT/dp"):
        pkg.RtaBatch(ctx, modes, [dict(runs[0], specularity=(0.0, 0.5, 0.5, 0.0), grad=True)],
                     tally="fp64-lds-runs", **kw)


def test_single_run_switches_tally_mode_around_spectral_locations(pkg, ctx):
    """bte_rta_execute keeps T, jx and jy in FP64 LDS sums (tally mode 2)
    unless spectral locations are set (then mode 0, device atomics, with the
    256-thread launch). One handle executed plain, with spectral locations
    and plain again follows the histories of fresh handles each time (equal
    event and crossing counts) and gives T, jx, jy equal to theirs up to the
    order of the FP64 sums."""
    W = pkg.workloads
    mat = W.material(200, seed=4)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    rect, kind, teq, _ = W.ribbon(300.0, 50.0, 30, 300.0, reservoirs=(300.5, 299.5))
    geom = pkg.Geometry(ctx, rect, kind, teq)
    kw = dict(thickness=1.0, T0=300.0, dt=5.0, maxtime=600.0, nparticles=40000, seed=5, sigma_bounce=2.0)

    def close(a, b):
        assert (a["events"], a["crossings"]) == (b["events"], b["crossings"])
        for key, off in (("T", 300.0), ("jx", 0.0), ("jy", 0.0)):
            scale = np.abs(b[key] - off).max()
            assert scale > 0
            np.testing.assert_allclose(a[key] - off, b[key] - off, rtol=0, atol=1e-10 * scale)

    fresh = pkg.rtamc2d(ctx, modes, geom, **kw)
    fresh_spec = pkg.rtamc2d(ctx, modes, geom, spectral={"boxes": [3, 10], "nomega": 20}, **kw)
    h = pkg.RtaRun(ctx, modes, geom, **kw)
    close(h.execute(), fresh)
    h.spectral([3, 10], 20)
    got = h.execute()
    close(got, fresh_spec)
    for a, b in zip(got["spectral"], fresh_spec["spectral"]):
        np.testing.assert_allclose(a["dT"], b["dT"], rtol=0, atol=1e-10 * np.abs(b["dT"]).max())
    h.spectral([], 0)
    close(h.execute(), fresh)
    h.close()


def test_rta_execute_switches_tally_modes_on_one_handle(pkg, ctx):
    """bte_rta_execute uses tally mode 2 while the cells fit and no spectral
    location is set, and mode 0 (256-thread workgroups) with one: a handle
    executed plain, with a spectral location and plain again gives each time
    the histories of a fresh handle (equal event and crossing counts) and T,
    jx, jy equal up to the order of the FP64 sums."""
    W = pkg.workloads
    mat = W.material(400, seed=6)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    rect, kind, teq, _ = W.ribbon(200.0, 50.0, 20, 300.0, reservoirs=(300.5, 299.5))
    geom = pkg.Geometry(ctx, rect, kind, teq)
    kw = dict(thickness=1.0, T0=300.0, dt=10.0, maxtime=300.0, nparticles=50000, seed=9, sigma_bounce=2.0)

    def fresh(spec):
        h = pkg.RtaRun(ctx, modes, geom, **kw)
        if spec:
            h.spectral([3], 50)
        o = h.execute()
        h.close()
        return o

    h = pkg.RtaRun(ctx, modes, geom, **kw)
    a = h.execute()
    h.spectral([3], 50)
    b = h.execute()
    h.spectral([], 1)
    c = h.execute()
    h.close()
    for got, ref in ((a, fresh(False)), (b, fresh(True)), (c, fresh(False))):
        assert (got["events"], got["crossings"]) == (ref["events"], ref["crossings"])
        for key, off in (("T", 300.0), ("jx", 0.0), ("jy", 0.0)):
            scale = np.abs(ref[key] - off).max()
            assert scale > 0
            np.testing.assert_allclose(got[key] - off, ref[key] - off, rtol=0, atol=1e-10 * scale)
