"""RTAMC2D steady branch (nonzero gradient) on the GPU: the bulk
relaxation-time flux in a cell periodic in x and y, and parity with the
oracle on a bar with reservoirs, diffuse walls and a gradient."""
import numpy as np
import pytest

import oracle_lib as orc
from stat_rules import assert_3sigma
from test_gpu_parity import zstats
from test_oracle_pins import onesample_z
from test_oracle_mc import bar, bulk_flux, inversion_symmetric, periodic_cell

pytestmark = pytest.mark.gpu


def test_steady_periodic_cell_gives_bulk_flux(pkg, ctx):
    m = inversion_symmetric(orc.synthetic_modes(40, 21))
    rect, kind, teq, per = periodic_cell()
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, periodic=per)
    g = (1.0e-3, 0.0)
    js = []
    for s in range(8):
        o = pkg.rta_steady(ctx, modes, geom, thickness=1.0, T0=300.0, gradient=g, nparticles=100000, seed=s,
                           sigma_bounce=5.0, max_generations=60)
        js.append((o["jx"][0], o["jy"][0]))
    js = np.array(js)
    j0 = np.asarray(bulk_flux(m, g))
    assert_3sigma(onesample_z(js, j0), "bulk flux against -kappa grad T")
    assert (js.std(0, ddof=1) / np.sqrt(len(js)) < 0.03 * np.abs(j0)).all()


def test_steady_bar_matches_oracle(pkg, ctx):
    W = pkg.workloads
    mat = W.material(300, seed=5)
    rect, kind, teq = bar(nbins=8, dx=10.0, H=30.0)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    kw = dict(thickness=1.0, T0=300.0, gradient=(-2.0e-3, 0.0), nparticles=20000, max_generations=40)
    R = 10
    g = [pkg.rta_steady(ctx, modes, geom, seed=100 + r, sigma_bounce=2.0, **kw) for r in range(R)]
    o = [orc.rta_steady_run(mat, rect, kind, teq, seed=200 + r, sigma=2.0, **kw) for r in range(R)]
    zs = [zstats(np.array([x[key][:8] for x in g]), np.array([x[key][:8] for x in o])) for key in ("T", "jx")]
    zs.append(zstats(np.array([x["jx"][:8].mean() for x in g]), np.array([x["jx"][:8].mean() for x in o])))
    assert_3sigma(zs, "steady bar vs oracle")
