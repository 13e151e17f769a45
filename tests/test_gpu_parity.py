"""GPU parity of the HIP path against the oracle on the same seeded inputs.

Monte Carlo results are compared statistically: R independent replicas on
each side give per-cell means and standard errors, and every compared cell
and a few scalar summaries of each run (spatial means, gradients, event
counts) form one family of z-scores that must pass the 3-sigma rule of
stat_rules.py (the count beyond 3 sigma within its binomial bound, the
largest within the Bonferroni bound). Cells of one run are correlated (a
replica with more flux has more flux everywhere), so no statistic pools z^2
over cells."""
import numpy as np
import pytest

import oracle_lib as orc
from stat_rules import assert_3sigma, welch_z

pytestmark = pytest.mark.gpu


def zstats(a, b):
    """a, b: (R, ...) replicas -> normal-equivalent z of the difference of
    means (Welch)"""
    return welch_z(a, b)


def _rta_setup(pkg, ctx, nbins=20, nmodes=600):
    W = pkg.workloads
    mat = W.material(nmodes, seed=3)
    rect, kind, teq, _ = W.ribbon(200.0, 100.0, nbins, 300.0, reservoirs=(300.5, 299.5))
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    return mat, rect, kind, teq, modes, geom


def test_rta_matches_oracle(pkg, ctx):
    mat, rect, kind, teq, modes, geom = _rta_setup(pkg, ctx)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=400.0)
    R, N = 16, 20000
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=100 + r, sigma_bounce=1.0, **kw) for r in range(R)]
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=200 + r, sigma=1.0, **kw) for r in range(R)]
    zs = []
    for key in ("T", "jx"):
        ga = np.array([x[key][[4, -1], :20] for x in g])
        oa = np.array([x[key][[4, -1], :20] for x in o])
        zs.append(zstats(ga, oa))
    summaries = {
        "jx_mean": lambda x: x["jx"][-1, :20].mean(),
        "T_gradient": lambda x: x["T"][-1, :10].mean() - x["T"][-1, 10:20].mean(),
        "T_early": lambda x: x["T"][4, :20].mean(),
        "events": lambda x: float(x["events"]),
        "crossings": lambda x: float(x["crossings"]),
    }
    for name, f in summaries.items():
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    assert_3sigma(zs, "RTAMC2D vs oracle")


def test_rta_per_mode_sigma_matches_oracle(pkg, ctx):
    """Per-mode bounce broadening (ModeTable.set_sigma, the reference's
    base_sigma with its low-outlier floor and pair width hypot(sigma_i,
    sigma_f), RTA_MC2D.cpp:566-618) on a narrow ribbon where most events are
    wall bounces: sigma grows with the mode's speed and a few modes sit far
    below the floor. 16 replicas per side."""
    W = pkg.workloads
    mat = W.material(2000, seed=12)
    sig = 0.3 + 1.5 * np.hypot(mat["vx"], mat["vy"]) / 8.0
    sig[::100] = 1e-5
    rect, kind, teq, _ = W.ribbon(200.0, 30.0, 20, 300.0, reservoirs=(300.5, 299.5))
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    modes.set_sigma(sig)
    geom = pkg.Geometry(ctx, rect, kind, teq)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=400.0)
    R, N = 16, 20000
    g = [pkg.rtamc2d(ctx, modes, geom, nparticles=N, seed=500 + r, sigma_bounce=1.0, **kw) for r in range(R)]
    o = [orc.rta_run(mat, rect, kind, teq, nparticles=N, seed=600 + r, sigma=1.0, mode_sigma=sig, **kw)
         for r in range(R)]
    zs = [zstats(np.array([x[key][[4, -1], :20] for x in g]), np.array([x[key][[4, -1], :20] for x in o]))
          for key in ("T", "jx")]
    for k in ("events", "crossings"):
        zs.append(zstats(np.array([float(x[k]) for x in g]), np.array([float(x[k]) for x in o])))
    assert_3sigma(zs, "RTAMC2D with per-mode sigma vs oracle")


def test_bert_matches_oracle(pkg, ctx):
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=300, seed=5)
    import scipy.linalg
    P = scipy.linalg.expm(5.0 * B)
    rect, kind, teq, tin = W.ribbon(100.0, 50.0, 10, 300.0, Tinit_first=301.0)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=5.0, maxtime=5.0 * 19, Efactor=20000, Tmax=301.0, Tmin=300.0)
    R = 12
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=10 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, tin, P=P, seed=20 + r, **kw) for r in range(R)]
    assert g[0]["nsteps"] == o[0]["nsteps"] == 20
    zs = []
    for q in (0, 1):
        ga = np.array([x["props"][[3, 10, 20], :, q] for x in g])
        oa = np.array([x["props"][[3, 10, 20], :, q] for x in o])
        zs.append(zstats(ga, oa))
    # centre of energy and mean flux at the last step
    xc = 0.5 * (rect[:, 0] + rect[:, 1])
    for name, f in {"centre": lambda x: (x["props"][-1, :, 0] * xc).sum() / x["props"][-1, :, 0].sum(),
                    "jx_mean": lambda x: x["props"][-1, :, 1].mean(),
                    "npart_last": lambda x: float(x["npart"][-1])}.items():
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    assert_3sigma(zs, "beRTAMC2D vs oracle")
    # P >= 0 and adiabatic walls conserve particles + pending wall counts
    # exactly (every particle here is positive)
    for x in g + o:
        tot = x["npart"] + x["pending"]
        assert (tot == tot[0]).all(), tot


def test_bert_given_initial_distribution_matches_oracle(pkg, ctx):
    """<init_distribution> (beyondRTA_MC2d.cpp:142-193, ref_distribution
    sampling.cpp:70-114): box 0 starts from a given signed distribution and
    box 3 from Tinit != Teq on top of it; both sets of particles are kept.
    The initial count, the energy and flux of every box at three steps and
    the particle count at the end against the oracle, 12 replicas per
    side."""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=240, seed=7)
    import scipy.linalg
    P = scipy.linalg.expm(4.0 * B)
    rect, kind, teq, tin = W.ribbon(100.0, 40.0, 8, 300.0)
    tin = tin.copy()
    tin[3] = 301.0
    n = len(mat["omega"])
    rng = np.random.default_rng(3)
    nb = orc.bose_einstein(mat["omega"], 302.0) - orc.bose_einstein(mat["omega"], 300.0)
    fd = 1e12 * 1.054571726e-34 * mat["omega"] * nb * np.where(rng.random(n) < 0.3, -1.0, 1.0)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=4.0, maxtime=4.0 * 19, Efactor=20000, Tmax=302.0, Tmin=300.0)
    R = 12
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=40 + r, init_dist={0: fd}, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, tin, P=P, seed=60 + r, init_dist={0: fd}, **kw) for r in range(R)]
    assert g[0]["npart"][0] > 1000 and np.abs(g[0]["props"][0, 0, 0]) > 0
    zs = [zstats(np.array([x["props"][[0, 5, 20], :8, q] for x in g]), np.array([x["props"][[0, 5, 20], :8, q] for x in o]))
          for q in (0, 1)]
    for f in (lambda x: float(x["npart"][0]), lambda x: float(x["npart"][-1])):
        zs.append(zstats(np.array([f(x) for x in g]), np.array([f(x) for x in o])))
    assert_3sigma(zs, "given initial distribution vs oracle")
    # a distribution of the wrong length is refused before any launch
    with pytest.raises(pkg.BteError, match="init_distribution"):
        pkg.BertRun(ctx, modes, geom, prop, seed=1, init_dist={0: fd[:-1]}, **kw)


def test_krylov_matches_oracle(pkg, ctx):
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=256, seed=7)
    for mode, mcode in (("reference", 0), ("textbook", 1)):
        Po = orc.build_P(B, 5.0, 30, mcode)
        Pg, rep = pkg.build_P(ctx, B, 5.0, m=30, precision="fp64", arnoldi=mode)
        assert np.abs(Pg - Po).max() < 1e-10, (mode, np.abs(Pg - Po).max())
    import scipy.linalg
    Pe = scipy.linalg.expm(5.0 * B)
    for prec, tol in (("fp64", 1e-11), ("fp32", 2e-5), ("bf16", 5e-5)):
        Pg, rep = pkg.build_P(ctx, B, 5.0, m=30, precision=prec, arnoldi="textbook")
        l1 = np.abs(Pg - Pe).sum(0).max()
        print(f"{prec}: max column L1 {l1:.3e}", flush=True)
        assert l1 < tol, (prec, l1)
        assert np.abs(Pg.sum(0) - 1).max() < 1e-12


@pytest.mark.parametrize("n", [16, 40, 257, 300, 520])
def test_krylov_bf16_sizes_off_the_product_tile(pkg, ctx, n):
    """The BF16 product runs on 256 x 256 tiles with the operator and basis
    padded by zero rows and columns: N = 300 (padded to 512; N not a multiple
    of 8, so the element-wise Gram-Schmidt kernels) and N = 520 (padded to
    768; the vector kernels) give the exact exponential within the BF16
    tolerance, columns summing to 1; so do N = 16, below the Krylov
    dimension m = 30 (the basis spans the space before m steps), N = 40 and
    N = 257, one past a tile."""
    import scipy.linalg
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=n, seed=11)
    Pe = scipy.linalg.expm(5.0 * B)
    assert np.abs(B).max() > 0
    Pg, rep = pkg.build_P(ctx, B, 5.0, m=30, precision="bf16", arnoldi="textbook")
    l1 = np.abs(Pg - Pe).sum(0).max()
    print(f"N={n} bf16: max column L1 {l1:.3e}, breakdowns {rep['breakdowns']}", flush=True)
    assert l1 < 5e-5, l1
    assert np.abs(Pg.sum(0) - 1).max() < 1e-12


def test_gpu_expm_matches_the_oracle(pkg, ctx):
    """bte_expm (Pade(13) with GPU GEMMs and LU) against the oracle's
    expm (pinned to scipy) on the synthetic operator, with squarings
    (norm above theta_13) and without, and on a diagonal matrix"""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=200, seed=3)
    for t in (0.1, 5.0):
        E, ms = pkg.expm(ctx, B, t)
        Eo = orc.expm(t * B)
        assert np.abs(E - Eo).max() < 1e-12 * max(1.0, np.abs(Eo).max()), (t, np.abs(E - Eo).max())
    d = np.linspace(-3.0, 2.0, 50)
    E, _ = pkg.expm(ctx, np.diag(d), 1.0)
    assert np.allclose(E, np.diag(np.exp(d)), rtol=1e-13, atol=1e-15)


def test_bert_reservoir_sources_match_oracle(pkg, ctx):
    """beRTAMC2D driven by reservoirs (Isothermal_walls, beyondRTA_MC2d.cpp:
    446-585): thousands of particles per step and source, so each source is
    split over several workgroups on the GPU."""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=200, seed=9)
    import scipy.linalg
    P = scipy.linalg.expm(2.0 * B)
    rect, kind, teq, _ = W.ribbon(60.0, 40.0, 6, 300.0, reservoirs=(301.0, 299.0))
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    prop = pkg.Propagator(ctx, P)
    kw = dict(thickness=1.0, dt=2.0, maxtime=2.0 * 24, Efactor=200000, Tmax=301.0, Tmin=299.0)
    R = 10
    g = [pkg.bertamc2d(ctx, modes, geom, prop, seed=70 + r, **kw) for r in range(R)]
    o = [orc.bert_run(mat, rect, kind, teq, teq, P=P, seed=90 + r, **kw) for r in range(R)]
    assert g[0]["npart"][-1] > 20000  # several thousand injected per step
    zs = [zstats(np.array([x["props"][[5, -1], :6, q] for x in g]), np.array([x["props"][[5, -1], :6, q] for x in o]))
          for q in (0, 1)]
    zs.append(zstats(np.array([float(x["npart"][-1]) for x in g]), np.array([float(x["npart"][-1]) for x in o])))
    assert_3sigma(zs, "reservoir-driven beRTAMC2D vs oracle")


def test_krylov_exponential_of_a_large_norm_hessenberg(pkg, ctx):
    """dt = 40 ps gives ||dt H||_1 = 68 on this operator (s = 8 halvings,
    2^s = 256 > m = 30), so the per-column exponential forms exp(H / 2^s)
    and squares it s times, while the oracle takes the exponential of its
    Arnoldi matrices by Pade scaling and squaring. The Krylov build has
    converged at this step: both agree with each other and with the full
    matrix exponential."""
    import scipy.linalg
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=128, seed=7)
    Po = orc.build_P(B, 40.0, 30, 1)
    Pg, rep = pkg.build_P(ctx, B, 40.0, m=30, precision="fp64", arnoldi="textbook")
    Pe = scipy.linalg.expm(40.0 * B)
    l1o, l1e = np.abs(Pg - Po).sum(0).max(), np.abs(Pg - Pe).sum(0).max()
    print(f"dt = 40 ps: max column L1 to the oracle {l1o:.3e}, to expm {l1e:.3e}", flush=True)
    assert l1o < 1e-10 and l1e < 1e-10
    assert np.abs(Pg.sum(0) - 1).max() < 1e-12
