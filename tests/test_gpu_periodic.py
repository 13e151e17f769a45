"""Periodic ghost boxes (geometry_2d.hpp:366-494) on the GPU: the exact
ballistic flux through a strip periodic in x, RTAMC2D and beRTAMC2D parity
with the oracle on that strip, and the front-end reading <Translate_to>."""
import numpy as np
import pytest

import oracle_lib as orc
from stat_rules import assert_3sigma
from test_gpu_parity import zstats
from test_oracle_mc import _bert_inputs, periodic_strip, tilted_y_material

pytestmark = pytest.mark.gpu


def test_periodic_ballistic_flux_is_exact(pkg, ctx):
    m = tilted_y_material()
    rect, kind, teq, per = periodic_strip()
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, periodic=per)
    out = pkg.rtamc2d(ctx, modes, geom, thickness=1.0, T0=300.0, dt=50.0, maxtime=1000.0, nparticles=100000,
                      seed=5, ballistic=True, sigma_bounce=5.0)
    vuc = m["V"] / m["lattz"]
    JB = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 0.0, 1.0, vuc)[2]
    JT = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, 0.0, -1.0, vuc)[2]
    np.testing.assert_allclose(out["jy"][-1, :5], (abs(JB) + abs(JT)) * 1e30, rtol=1e-9)
    assert np.abs(out["jy"][-1, 5:7]).max() == 0.0


def test_periodic_rta_matches_oracle(pkg, ctx):
    W = pkg.workloads
    mat = W.material(400, seed=8)
    rect, kind, teq, per = periodic_strip(width=30.0, nrows=6, dy=15.0)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, periodic=per)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=300.0)
    R, N = 12, 20000
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=300 + r, sigma_bounce=2.0, **kw) for r in range(R)]
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=400 + r, sigma=2.0, periodic=per, **kw)
         for r in range(R)]
    zs = [zstats(np.array([x[key][[3, -1], :6] for x in g]), np.array([x[key][[3, -1], :6] for x in o]))
          for key in ("T", "jy")]
    for name, f in {"jy_mean": lambda x: x["jy"][-1, :6].mean(), "events": lambda x: float(x["events"]),
                    "crossings": lambda x: float(x["crossings"])}.items():
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    assert_3sigma(zs, "periodic RTAMC2D vs oracle")


def test_periodic_bert_matches_oracle(pkg, ctx):
    m, B, P = _bert_inputs(48, seed=6)
    rect, kind, teq, per = periodic_strip(TB=300.0, TT=300.0)
    rect, kind, teq = rect[:7], kind[:7], teq[:7]  # closed strip: adiabatic walls below and above
    tin = teq.copy()
    tin[0] = 310.0
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin, periodic=per)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=5.0, maxtime=60.0, Efactor=5000, Tmax=310.0, Tmin=300.0)
    R = 12
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=30 + r, **kw) for r in range(R)]
    o = [orc.bert_run(m, rect, kind, teq, tin, P=P, seed=50 + r, periodic=per, **kw) for r in range(R)]
    for x in g:
        tot = x["npart"] + x["pending"]
        assert (tot == tot[0]).all()
    z = zstats(np.array([x["props"][[2, -1], :5, 0] for x in g]), np.array([x["props"][[2, -1], :5, 0] for x in o]))
    assert_3sigma(z, "periodic beRTAMC2D vs oracle")
