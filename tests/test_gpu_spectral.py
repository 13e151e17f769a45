"""RTAMC2D spectral decomposition on the GPU (per-mode tallies of the
spectral locations, Gamma broadening as one matrix product afterwards)
against the oracle, which broadens every segment as the reference does
(RTA_MC2D.cpp:718-766), and the executable's spectral CSV files."""
import numpy as np
import pytest
import scipy.stats

import oracle_lib as orc
from stat_rules import assert_3sigma
from test_gpu_parity import zstats
from test_gpu_frontends import run, write_ribbon
from test_oracle_spectral import one_frequency_material, ribbon4

pytestmark = pytest.mark.gpu


def _modes(pkg, ctx, mat):
    return pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])


def test_single_frequency_spectra_on_the_gpu(pkg, ctx):
    mat = one_frequency_material()
    rect, kind, teq = ribbon4()
    K, sigma, T0 = 40, 0.3, 300.0
    r = pkg.rtamc2d(ctx, _modes(pkg, ctx, mat), pkg.Geometry(ctx, rect, kind, teq), thickness=1.0, T0=T0, dt=10.0,
                    maxtime=200.0, nparticles=50000, seed=5, sigma_bounce=sigma,
                    spectral={"boxes": [1, 2], "nomega": K})
    grid = np.linspace(0.0, 5.0, K) + 0.5 * 5.0 / K
    w = scipy.stats.gamma.pdf(grid, a=25.0 / sigma, scale=sigma / 5.0)
    cv = orc.calc_cv(mat["omega"], mat["V"], T0)
    for k, box in enumerate((1, 2)):
        sp = r["spectral"][k]
        assert np.allclose(sp["omega"], grid, rtol=1e-14)
        dT = r["T"][:, box] - T0
        assert np.abs(dT).max() > 1e-4
        assert np.allclose(sp["dT"], np.outer(dT, w), rtol=1e-9, atol=1e-11 * np.abs(dT).max() * w.max())
        jx = r["jx"][:, box]
        assert np.allclose(sp["jx"], np.outer(jx, w), rtol=1e-9, atol=1e-11 * np.abs(jx).max() * w.max())
        assert np.allclose(sp["fd"][:, 0], dT * cv / (1e12 * 1.054571726e-34 * 5.0), rtol=1e-9,
                           atol=1e-11 * np.abs(sp["fd"]).max())


def test_spectral_decomposition_matches_the_oracle(pkg, ctx):
    mat = pkg.workloads.material(300, seed=6)
    rect, kind, teq, _ = pkg.workloads.ribbon(200.0, 60.0, 8, 300.0, reservoirs=(300.5, 299.5))
    spec = {"boxes": [1, 6], "nomega": 30}
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=400.0)
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, rect, kind, teq)
    R, N = 12, 20000
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=10 + r, sigma_bounce=1.0, spectral=spec, **kw)
         for r in range(R)]
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=40 + r, sigma=1.0, spectral=spec, **kw)
         for r in range(R)]
    assert np.allclose(g[0]["spectral"][0]["omega"], o[0]["spectral"][0]["omega"], rtol=1e-14)
    zs = []
    for k in range(2):
        for key in ("dT", "jx"):
            # rows 4 and last; ticks summed in bands of 5
            ga = np.array([x["spectral"][k][key][[4, -1]].reshape(2, 6, 5).sum(-1) for x in g])
            oa = np.array([x["spectral"][k][key][[4, -1]].reshape(2, 6, 5).sum(-1) for x in o])
            zs.append(zstats(ga, oa))
        zs.append(zstats(np.array([x["spectral"][k]["fd"][-1] for x in g]),
                         np.array([x["spectral"][k]["fd"][-1] for x in o])))
    assert_3sigma(zs, "spectral tallies vs oracle")


def test_rtamc2d_writes_spectral_files(pkg, ctx, tmp_path):
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(60.0, 30.0, 6, 300.0, reservoirs=(300.5, 299.5))
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, teq)
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d>\n<geometry file="g.xml"/>\n<time dt="20" maxtime="200"/>\n'
        '<material name="m" modetable="synthetic" nmodes="400" seed="2" thickness="1" T0="300" sigma_bounce="2.0"/>\n'
        '<particles N="20000" seed="5"/>\n<spectral>\n  <resolution ticks="25"/>\n  <location bin="2"/>\n'
        '  <location bin="4"/>\n</spectral>\n</RTA_MC2d>\n')
    run([pkg.app_path("RTAMC2D"), "in.xml", "8"], tmp_path)
    mat = W.material(400, seed=2)
    o = pkg.rtamc2d(ctx, _modes(pkg, ctx, mat), pkg.Geometry(ctx, rect, kind, teq), thickness=1.0, T0=300.0,
                    dt=20.0, maxtime=200.0, nparticles=20000, seed=5, sigma_bounce=2.0,
                    spectral={"boxes": [2, 4], "nomega": 25})
    for k, loc in enumerate((2, 4)):
        for key, name in (("dT", "deltaT_omega"), ("jx", "jx_omega"), ("jy", "jy_omega")):
            a = np.loadtxt(tmp_path / f"{name}_{loc}_300K_run_0.csv", delimiter=",")
            # ntime = 11 rows: the tick row, then row i at time (i + 0.5) dt
            assert a.shape == (11, 26)
            assert a[0, 0] == -1.0 and np.allclose(a[0, 1:], o["spectral"][k]["omega"])
            assert np.allclose(a[1:, 0], (np.arange(1, 11) + 0.5) * 20.0)
            ref = o["spectral"][k][key]
            assert np.allclose(a[1:, 1:], ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
        lines = (tmp_path / f"fd_q_{loc}_300K_run_0.csv").read_text().splitlines()
        assert lines[0] == "# 0.000000" and lines[1] == "# 0.000000" and lines[2].startswith("# ")
        fd = np.array([float(lines[2][2:])] + [float(x) for x in lines[3:]])
        assert np.allclose(fd, o["spectral"][k]["fd"][:, 0], rtol=1e-5)
