"""GPU parity at the benchmark configurations' own inputs: the mode tables,
geometries, time grids and operators of configurations 0, 2, 3 and 4 (see
bte-barna_amd/workloads.py), at particle counts the oracle finishes in
seconds, against the oracle on identical inputs.

Monte Carlo outputs (binned T, flux and dT/dp at several times) come from R
independent replicas on each side; all z-scores of one configuration form
one family under the 3-sigma rule of stat_rules.py. The Krylov build is
deterministic and is compared column by column with stated tolerances."""
import numpy as np
import pytest

import oracle_lib as orc
from stat_rules import assert_3sigma, exact_z, welch_z

pytestmark = pytest.mark.gpu


def _modes(pkg, ctx, mat):
    return pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])


def test_config0_rtamc2d_matches_oracle(pkg, ctx):
    """Configuration 0 (3000 modes, 1 um x 200 nm, 100 bins, walls at
    300 +- 0.5 K, dt 25 ps, 5 ns) at 1e5 particles per replica, 12 replicas
    per side, through the single-run path (FP64 LDS tallies): T and jx of
    every bin at 4 times, plus event and crossing counts."""
    W = pkg.workloads
    c = W.config0(nparticles=100_000)
    mat = c["material"]
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, c["rect"], c["kind"], c["Teq"])
    kw = dict(thickness=c["thickness"], T0=c["T0"], dt=c["dt"], maxtime=c["maxtime"], nparticles=c["nparticles"])
    R = 12
    g = [pkg.rtamc2d(ctx, modes, geom, seed=1000 + r, sigma_bounce=c["sigma_bounce"], **kw) for r in range(R)]
    o = [orc.rta_run(mat, c["rect"], c["kind"], c["Teq"], seed=2000 + r, sigma=c["sigma_bounce"], **kw)
         for r in range(R)]
    rows = [3, 19, 79, 199]
    zs = [welch_z(np.array([x[k][rows, :100] for x in g]), np.array([x[k][rows, :100] for x in o]))
          for k in ("T", "jx")]
    for k in ("events", "crossings"):
        zs.append(welch_z(np.array([float(x[k]) for x in g]), np.array([float(x[k]) for x in o])))
    print(assert_3sigma(zs, "configuration 0"))


def test_config2_krylov_precisions_match_the_fp64_oracle(pkg, ctx):
    """Configuration 2 (N = 8192, m = 30, dt = 5 ps, synthetic Omega): the
    GPU block-Krylov build in FP64, FP32 and BF16 (hi + lo) against the
    oracle's per-column FP64 Arnoldi (textbook projections) on 32 columns;
    largest column L1 distance within 1e-12, 1e-6 and 1e-5. Every column of
    each build sums to 1 within 1e-12 after the FP64 energy projection, and
    no entry is below minus the same tolerance (the exact propagator is
    nonnegative)."""
    W = pkg.workloads
    mat, B = W.config1_inputs(nmodes=8192, seed=1)
    blocks = [(0, 16), (4096, 4112)]
    ref = [orc.build_P_cols(B, 5.0, 30, 1, a, b) for a, b in blocks]
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6), ("bf16", 1e-5)):
        P, rep = pkg.build_P(ctx, B, 5.0, m=30, precision=prec, arnoldi="textbook")
        l1 = max(float(np.abs(P[:, a:b] - r).sum(0).max()) for (a, b), r in zip(blocks, ref))
        drift = float(np.abs(P.sum(0) - 1.0).max())
        print(f"{prec}: max column L1 {l1:.3e}, column-sum drift {drift:.1e}, min entry {P.min():.2e}, "
              f"{rep['build_ms']:.1f} ms", flush=True)
        assert l1 <= tol, (prec, l1)
        assert drift <= 1e-12, (prec, drift)
        assert P.min() >= -tol, (prec, P.min())
        del P


def test_config3_extreme_runs_match_oracle(pkg, ctx):
    """Configuration 3's narrowest and coldest run (50 nm, 100 K), its middle
    run and its widest and hottest run (2 um, 500 K): each with its own T0
    (Cv and the BE_derivative sampler at T0) and scattering-rate factor
    T0 / 300 K, 12 replicas of 3e4 particles per run and side. The GPU runs
    all 36 in one batched launch; T and jx of every bin at 3 times and the
    event counts per run form one family."""
    W = pkg.workloads
    c = W.config3()
    mat = c["material"]
    modes = _modes(pkg, ctx, mat)
    pick = [0, 31, 63]
    R, N = 12, 30000
    runs, geoms = [], {}
    for i in pick:
        rd = c["runs"][i]
        geoms[i] = pkg.Geometry(ctx, rd["rect"], rd["kind"], rd["Teq"])
        for r in range(R):
            runs.append(dict(geom=geoms[i], T0=rd["T0"], nparticles=N, seed=5000 + 100 * i + r, wscale=rd["wscale"]))
    b = pkg.RtaBatch(ctx, modes, runs, thickness=c["thickness"], dt=c["dt"], maxtime=c["maxtime"],
                     sigma_bounce=c["sigma_bounce"])
    res, rep = b.execute()
    b.close()
    rows = [19, 79, 199]
    zs = []
    ev_o, ev_var = 0.0, 0.0
    for j, i in enumerate(pick):
        rd = c["runs"][i]
        g = res[j * R:(j + 1) * R]
        o = [orc.rta_run(mat, rd["rect"], rd["kind"], rd["Teq"], thickness=c["thickness"], T0=rd["T0"], dt=c["dt"],
                         maxtime=c["maxtime"], nparticles=N, seed=7000 + 100 * i + r, sigma=c["sigma_bounce"],
                         wscale=rd["wscale"]) for r in range(R)]
        for k in ("T", "jx"):
            zs.append(welch_z(np.array([x[k][rows, :100] for x in g]), np.array([x[k][rows, :100] for x in o])))
        ev = np.array([float(x["events"]) for x in o])
        ev_o += ev.sum()
        ev_var += R * ev.var(ddof=1)
    # the batch reports its total event count: against the oracle's total,
    # each side with the variance of the oracle's runs
    zs.append(exact_z(float(rep["events"]), ev_o, np.sqrt(2.0 * ev_var)))
    print(assert_3sigma(zs, "configuration 3 runs 0, 31, 63"))


def test_config4_specular_gradient_matches_oracle(pkg, ctx):
    """Configuration 4 (1 um x 100 nm, 128 bins x 64 time bins, 5000 modes in
    v_y mirror pairs, p_bottom = 0.7, p_top = 0.3) with its FP32 LDS tallies
    and FP64 final sums, 12 replicas of 5e4 particles per side: T and jx
    of every bin at 3 times and the score-function dT/dp of both edges at
    2 times against the oracle's specular reflection and score tallies."""
    W = pkg.workloads
    c = W.config4(nparticles=50_000)
    mat = c["material"]
    modes = _modes(pkg, ctx, mat)
    geom = pkg.Geometry(ctx, c["rect"], c["kind"], c["Teq"])
    R, N = 12, c["nparticles"]
    # the FP32 LDS tallies hold one run's 128 x 64 cells: one batch per replica
    g, ev_g = [], 0
    for r in range(R):
        run = dict(geom=geom, T0=c["T0"], nparticles=N, seed=300 + r, specularity=c["specularity"], grad=True)
        b = pkg.RtaBatch(ctx, modes, [run], thickness=c["thickness"], dt=c["dt"], maxtime=c["maxtime"],
                         sigma_bounce=c["sigma_bounce"], tally="fp32-lds")
        (x,), rep = b.execute(grad=True)
        b.close()
        x["events"] = rep["events"]
        g.append(x)
    o = [orc.rta_run(mat, c["rect"], c["kind"], c["Teq"], thickness=c["thickness"], T0=c["T0"], dt=c["dt"],
                     maxtime=c["maxtime"], nparticles=N, seed=900 + r, sigma=c["sigma_bounce"],
                     specularity=c["specularity"], grad=True) for r in range(R)]
    rows = [7, 31, 63]
    zs = [welch_z(np.array([x[k][rows, :128] for x in g]), np.array([x[k][rows, :128] for x in o]))
          for k in ("T", "jx")]
    for side in (1, 2):
        zs.append(welch_z(np.array([x["dTdp"][side][[31, 63], :128] for x in g]),
                          np.array([x["dTdp"][side][[31, 63], :128] for x in o])))
    zs.append(welch_z(np.array([float(x["events"]) for x in g]), np.array([float(x["events"]) for x in o])))
    print(assert_3sigma(zs, "configuration 4"))
