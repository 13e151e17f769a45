"""Convex polygon boxes on the GPU (geometry_2d.cpp:27-522 on general convex
hulls; RTA_MC2D.cpp:990-1386 and beyondRTA_MC2d.cpp:281-970 on their sides):
the library's contacts against the oracle's, and RTAMC2D and beRTAMC2D on
polygon geometries (slanted shared sides, slanted walls with their own
diffuse-bounce and emission normals, reservoirs, a pulse in a polygon box,
gradient sources at polygon centroids) against the oracle under the 3-sigma
rule. A rectangle geometry given as polygons runs the polygon kernels and
must agree with the oracle's rectangles too."""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib as orc
from stat_rules import assert_3sigma
from test_gpu_parity import zstats

pytestmark = pytest.mark.gpu


def zigzag_ribbon(n, L, H, slant, wave, reservoirs=False, res_w=10.0):
    """n quadrilateral cells along x: cell i has bottom corners x_i, x_i+1
    at y = 0 and top corners shifted by +-slant in x and +-wave in y
    (alternating), so neighbours share slanted sides and the top wall is a
    zigzag; the two end sides are vertical. With reservoirs, rectangles of
    width res_w at both ends (boxes n and n + 1)."""
    xb = np.linspace(0.0, L, n + 1)
    s = slant * (-1.0) ** np.arange(n + 1)
    s[0] = s[-1] = 0.0
    h = H + wave * (-1.0) ** np.arange(n + 1)
    polys = [np.array([[xb[i], 0.0], [xb[i + 1], 0.0], [xb[i + 1] + s[i + 1], h[i + 1]], [xb[i] + s[i], h[i]]])
             for i in range(n)]
    if reservoirs:
        polys.append(np.array([[-res_w, 0.0], [0.0, 0.0], [0.0, h[0]], [-res_w, h[0]]]))
        polys.append(np.array([[L, 0.0], [L + res_w, 0.0], [L + res_w, h[-1]], [L, h[-1]]]))
    return polys


def _modes(pkg, ctx, mat):
    return pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])


def test_polygon_contacts_match_the_oracle(pkg, ctx):
    polys = zigzag_ribbon(8, 80.0, 30.0, 3.0, 4.0, reservoirs=True)
    r = 3.0
    hexes = [np.array([[cx + r * np.cos(a), cy + r * np.sin(a)] for a in np.arange(6) * np.pi / 3])
             for cx, cy in ((0.0, 0.0), (1.5 * r, np.sqrt(3) / 2 * r), (1.5 * r, -np.sqrt(3) / 2 * r))]
    for ps in (polys, hexes):
        n = len(ps)
        geom = pkg.Geometry(ctx, None, np.zeros(n, np.int32), np.full(n, 300.0), polys=ps)
        nc, nbr, geo = geom.contacts()
        onc, onbr, ogeo = orc.contacts(None, polys=ps)
        assert np.array_equal(nc, onc)
        for i in range(n):
            got = {int(nbr[i, k]): geo[i, k] for k in range(nc[i])}
            ref = {int(onbr[i, k]): ogeo[i, k] for k in range(onc[i])}
            assert got.keys() == ref.keys()
            for j in got:
                assert np.allclose(got[j][0:2], ref[j][0:2], atol=1e-12)
                seg, oseg = got[j][2:6].reshape(2, 2), ref[j][2:6].reshape(2, 2)
                assert np.allclose(np.sort(seg, 0), np.sort(oseg, 0), atol=1e-9)
    with pytest.raises(pkg.BteError, match="overlapping"):
        pkg.Geometry(ctx, None, np.zeros(2, np.int32), np.full(2, 300.0), polys=[hexes[0], hexes[0] + 1.0])


def _rta_compare(pkg, ctx, mat, polys, kind, teq, tag, R=12, N=20000):
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, None, kind, teq, polys=polys)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=400.0)
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=100 + r, sigma_bounce=1.0, **kw) for r in range(R)]
    rect = orc.poly_rect(polys)
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=200 + r, sigma=1.0, polys=polys, **kw)
         for r in range(R)]
    nreg = int((kind == 0).sum())
    zs = []
    for key in ("T", "jx", "jy"):
        zs.append(zstats(np.array([x[key][[4, -1], :nreg] for x in g]), np.array([x[key][[4, -1], :nreg] for x in o])))
    for k in ("events", "crossings"):
        zs.append(zstats(np.array([float(x[k]) for x in g]), np.array([float(x[k]) for x in o])))
    print(assert_3sigma(zs, tag))
    return g


def test_rta_on_zigzag_polygons_matches_oracle(pkg, ctx):
    mat = pkg.workloads.material(600, seed=3)
    polys = zigzag_ribbon(12, 200.0, 80.0, 6.0, 9.0, reservoirs=True)
    kind = np.array([0] * 12 + [1, 1], np.int32)
    teq = np.array([300.0] * 12 + [300.5, 299.5])
    g = _rta_compare(pkg, ctx, mat, polys, kind, teq, "RTAMC2D zigzag polygons vs oracle")
    # heat flows from the hot reservoir (left) to the cold one
    assert np.mean([x["jx"][-1, :12].mean() for x in g]) > 0


def test_rta_rectangles_as_polygons_match_oracle(pkg, ctx):
    mat = pkg.workloads.material(600, seed=3)
    rect, kind, teq, _ = pkg.workloads.ribbon(200.0, 100.0, 10, 300.0, reservoirs=(300.5, 299.5))
    polys = [np.array([[r[0], r[2]], [r[0], r[3]], [r[1], r[3]], [r[1], r[2]]]) for r in rect]
    _rta_compare(pkg, ctx, mat, polys, kind, teq, "RTAMC2D rectangles through the polygon kernels vs oracle")


def test_bert_on_zigzag_polygons_matches_oracle(pkg, ctx):
    """adiabatic zigzag walls (emission along slanted normals), slanted
    shared sides and an initial pulse in the first polygon box"""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=300, seed=5)
    P = scipy.linalg.expm(5.0 * B)
    polys = zigzag_ribbon(10, 100.0, 40.0, 3.0, 5.0)
    n = len(polys)
    kind = np.zeros(n, np.int32)
    teq = np.full(n, 300.0)
    tin = teq.copy()
    tin[0] = 301.0
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, None, kind, teq, tin, polys=polys)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=5.0, maxtime=5.0 * 19, Efactor=20000, Tmax=301.0, Tmin=300.0)
    R = 12
    rect = orc.poly_rect(polys)
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=10 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, tin, P=P, seed=20 + r, polys=polys, **kw) for r in range(R)]
    assert g[0]["nsteps"] == o[0]["nsteps"] == 20
    zs = []
    for q in (0, 1, 2):
        zs.append(zstats(np.array([x["props"][[3, 10, 20], :, q] for x in g]),
                         np.array([x["props"][[3, 10, 20], :, q] for x in o])))
    zs.append(zstats(np.array([float(x["npart"][-1]) for x in g]), np.array([float(x["npart"][-1]) for x in o])))
    print(assert_3sigma(zs, "beRTAMC2D zigzag polygons vs oracle"))
    # P >= 0 and adiabatic walls: particles + pending wall counts conserved
    for x in g + o:
        tot = x["npart"] + x["pending"]
        assert (tot == tot[0]).all(), tot


def test_bert_polygon_reservoirs_and_gradient_sources_match_oracle(pkg, ctx):
    """reservoir injection through vertical end sides and gradient sources
    from calculate_gradientT at the polygon centroids (uniform points of a
    polygon box by bounding-box rejection)"""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=200, seed=9)
    P = scipy.linalg.expm(2.0 * B)
    polys = zigzag_ribbon(6, 60.0, 40.0, 4.0, 6.0, reservoirs=True)
    kind = np.array([0] * 6 + [1, 1], np.int32)
    cx = np.array([p[:, 0].mean() for p in polys])
    teq = np.concatenate([300.0 + 0.05 * cx[:6], [301.0, 299.0]])
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, None, kind, teq, polys=polys)
    assert np.allclose(geom.gradient_T(), orc.gradient_T(None, kind, teq, polys=polys), rtol=1e-10, atol=1e-15)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 20, Efactor=50000, Tmax=303.0, Tmin=299.0)
    R = 10
    rect = orc.poly_rect(polys)
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=70 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, teq, P=P, seed=90 + r, polys=polys, **kw) for r in range(R)]
    assert g[0]["npart"][-1] > 10000
    zs = [zstats(np.array([x["props"][[5, -1], :6, q] for x in g]), np.array([x["props"][[5, -1], :6, q] for x in o]))
          for q in (0, 1, 2)]
    zs.append(zstats(np.array([float(x["npart"][-1]) for x in g]), np.array([float(x["npart"][-1]) for x in o])))
    print(assert_3sigma(zs, "polygon reservoirs and gradient sources vs oracle"))


def test_rta_steady_on_zigzag_polygons_matches_oracle(pkg, ctx):
    """steady branch with a gradient (RTA_MC2D.cpp:1406-1799, 2544-2842):
    gradient generators and re-emission at uniform points of polygon boxes,
    flights through slanted sides, diffuse zigzag walls"""
    mat = pkg.workloads.material(300, seed=5)
    polys = zigzag_ribbon(8, 80.0, 30.0, 3.0, 4.0, reservoirs=True)
    kind = np.array([0] * 8 + [1, 1], np.int32)
    teq = np.array([300.0] * 8 + [300.5, 299.5])
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, None, kind, teq, polys=polys)
    kw = dict(thickness=1.0, T0=300.0, gradient=(-2.0e-3, 0.0), nparticles=20000, max_generations=40)
    R = 10
    rect = orc.poly_rect(polys)
    g = [pkg.rta_steady(ctx, modes, geom, seed=100 + r, sigma_bounce=2.0, **kw) for r in range(R)]
    o = [orc.rta_steady_run(mat, rect, kind, teq, seed=200 + r, sigma=2.0, polys=polys, **kw) for r in range(R)]
    zs = [zstats(np.array([x[key][:8] for x in g]), np.array([x[key][:8] for x in o])) for key in ("T", "jx", "jy")]
    zs.append(zstats(np.array([x["jx"][:8].mean() for x in g]), np.array([x["jx"][:8].mean() for x in o])))
    print(assert_3sigma(zs, "steady zigzag polygons vs oracle"))


def sheared_strip(width=20.0, nrows=5, dy=10.0, shear=0.5, TB=300.5, TT=299.5):
    """periodic_strip sheared by x -> x + shear * y: a column of
    parallelograms periodic in x through parallelogram ghost boxes on both
    sides (each translates onto the other), reservoirs below and above"""
    H = nrows * dy

    def X(x, y):
        return x + shear * y

    polys = [np.array([[X(0, a), a], [X(width, a), a], [X(width, b), b], [X(0, b), b]])
             for a, b in ((k * dy, (k + 1) * dy) for k in range(nrows))]
    polys.append(np.array([[X(-2, 0), 0], [X(0, 0), 0], [X(0, H), H], [X(-2, H), H]]))
    polys.append(np.array([[X(width, 0), 0], [X(width + 2, 0), 0], [X(width + 2, H), H], [X(width, H), H]]))
    polys.append(np.array([[-2.0, -2.0], [width + 2.0, -2.0], [width + 2.0, 0.0], [-2.0, 0.0]]))
    polys.append(np.array([[X(-2, H), H], [X(width + 2, H), H], [X(width + 2, H), H + 2.0], [X(-2, H), H + 2.0]]))
    kind = np.array([0] * nrows + [2, 2, 1, 1], np.int32)
    teq = np.array([300.0] * (nrows + 2) + [TB, TT])
    left, right = nrows, nrows + 1
    return polys, kind, teq, {left: (right, width, 0.0), right: (left, -width, 0.0)}


def test_sheared_periodic_strip_ballistic_flux_is_exact(pkg, ctx):
    """translate on polygon ghost boxes: ballistic particles from the bottom
    reservoir wrap around the sheared periodic direction until they reach
    the top, so every row carries the flux (|J_B| + |J_T|) 1e30 (each
    particle crosses each row once, dy / vy in it)"""
    from test_oracle_mc import tilted_y_material
    m = tilted_y_material()
    polys, kind, teq, per = sheared_strip()
    modes = _modes(pkg, ctx, m)
    geom = pkg.Geometry(ctx, None, kind, teq, periodic=per, polys=polys)
    out = pkg.rtamc2d(ctx, modes, geom, thickness=1.0, T0=300.0, dt=50.0, maxtime=1000.0, nparticles=100000,
                      seed=5, ballistic=True, sigma_bounce=5.0)
    vuc = m["V"] / m["lattz"]
    JB = orc.wall_dist(m["omega"], m["vx"], m["vy"], 300.5, 300.0, 0.0, 1.0, vuc)[2]
    JT = orc.wall_dist(m["omega"], m["vx"], m["vy"], 299.5, 300.0, 0.0, -1.0, vuc)[2]
    np.testing.assert_allclose(out["jy"][-1, :5], (abs(JB) + abs(JT)) * 1e30, rtol=1e-8)
    assert np.abs(out["jy"][-1, 5:7]).max() == 0.0
    assert out["events"] > 0


def test_sheared_periodic_strip_matches_oracle(pkg, ctx):
    mat = pkg.workloads.material(400, seed=8)
    polys, kind, teq, per = sheared_strip(width=30.0, nrows=6, dy=15.0)
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, None, kind, teq, periodic=per, polys=polys)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=300.0)
    R, N = 12, 20000
    rect = orc.poly_rect(polys)
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=300 + r, sigma_bounce=2.0, **kw) for r in range(R)]
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=400 + r, sigma=2.0, periodic=per, polys=polys, **kw)
         for r in range(R)]
    zs = [zstats(np.array([x[key][[3, -1], :6] for x in g]), np.array([x[key][[3, -1], :6] for x in o]))
          for key in ("T", "jx", "jy")]
    for k in ("events", "crossings"):
        zs.append(zstats(np.array([float(x[k]) for x in g]), np.array([float(x[k]) for x in o])))
    print(assert_3sigma(zs, "sheared periodic strip vs oracle"))


def test_polygon_geometry_errors(pkg, ctx):
    """a vertex between two hull corners (the reference would keep part of
    that side as a border), a degenerate box and too many sides are refused"""
    k1, t1 = np.zeros(1, np.int32), np.full(1, 300.0)
    with pytest.raises(pkg.BteError, match="vertex lies on a side"):
        pkg.Geometry(ctx, None, k1, t1, polys=[np.array([[0.0, 0.0], [5.0, 0.0], [10.0, 0.0], [10.0, 5.0], [0.0, 5.0]])])
    with pytest.raises(pkg.BteError, match="degenerate"):
        pkg.Geometry(ctx, None, k1, t1, polys=[np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0]])])
    a = 2 * np.pi * np.arange(20) / 20
    with pytest.raises(pkg.BteError, match="more than 16 sides"):
        pkg.Geometry(ctx, None, k1, t1, polys=[np.stack([np.cos(a), np.sin(a)], 1)])
