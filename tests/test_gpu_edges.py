"""Boundary cases of the GPU path: empty populations, a single closed box,
one-particle and two-mode runs, and argument errors that the library turns
into BteError before any launch. The oracle runs the same empty and
single-box cases (beyondRTA_MC2d.cpp and RTA_MC2D.cpp read the same inputs:
a geometry at equilibrium generates no particle, a closed box keeps its
energy)."""
import numpy as np
import pytest

import oracle_lib as orc
from test_oracle_mc import _bert_inputs, bar

pytestmark = pytest.mark.gpu


def _modes(pkg, ctx, m):
    return pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])


def test_bert_at_equilibrium_has_no_particles(pkg, ctx):
    m, B, P = _bert_inputs(40, seed=3)
    rect, kind, teq, tin = pkg.workloads.ribbon(60.0, 30.0, 6, 300.0)
    assert (tin == teq).all()
    kw = dict(thickness=1.0, dt=5.0, maxtime=45.0, Efactor=100000, Tmax=301.0, Tmin=300.0, seed=3)
    g = pkg.bertamc2d(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq, tin), pkg.Propagator(ctx, P), **kw)
    o = orc.bert_run(m, rect, kind, teq, tin, P=P, **kw)
    assert g["nsteps"] == o["nsteps"] == 10
    for r in (g, o):
        assert (r["npart"] == 0).all() and (r["pending"] == 0).all()
        assert not r["props"].any()
    assert g["particle_steps"] == 0
    assert g["Eeff"] == pytest.approx(o["Eeff"], rel=1e-12)


def test_bert_single_closed_box_keeps_its_energy(pkg, ctx):
    """one box, adiabatic diffuse walls on every side: each wall hit is
    re-emitted in the same step, so particles plus pending wall counts stay
    fixed and the energy density follows the particle count"""
    m, B, P = _bert_inputs(40, seed=5)
    rect = np.array([[0.0, 20.0, 0.0, 20.0]])
    kind = np.zeros(1, np.int32)
    teq, tin = np.array([300.0]), np.array([302.0])
    kw = dict(thickness=1.0, dt=5.0, maxtime=95.0, Efactor=200000, Tmax=302.0, Tmin=300.0, seed=9)
    g = pkg.bertamc2d(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq, tin), pkg.Propagator(ctx, P), **kw)
    o = orc.bert_run(m, rect, kind, teq, tin, P=P, **kw)
    assert g["npart"][0] == o["npart"][0] > 0
    for r in (g, o):
        tot = r["npart"] + r["pending"]
        assert (tot == tot[0]).all()
        count = r["props"][:, 0, 0] * 400.0 / (r["Eeff"] * 1e27)
        np.testing.assert_allclose(count, r["npart"], rtol=1e-9)


def test_rta_with_one_particle(pkg, ctx):
    m = pkg.workloads.material(200, seed=2)
    rect, kind, teq = bar(nbins=4, TL=301.0, TR=299.0)
    r = pkg.rtamc2d(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq), thickness=1.0, T0=300.0, dt=10.0,
                    maxtime=100.0, nparticles=1, seed=4, sigma_bounce=2.0)
    assert r["events"] >= 1
    assert np.isfinite(r["T"]).all() and np.isfinite(r["jx"]).all()


def test_rta_two_mode_material(pkg, ctx):
    """the smallest table with both reservoirs emitting: one mode and its
    mirror image through the origin (alias tables of two entries); GPU and
    oracle agree on Eeff"""
    m = orc.synthetic_modes(2, 7, mirror=True)
    m["vx"] = np.array([1.0, -1.0]) * abs(m["vx"][0])
    m["vy"] = np.array([1.0, -1.0]) * abs(m["vy"][0])
    m["V"], m["lattz"] = 20.0, 1.0
    rect, kind, teq = bar(nbins=4, TL=301.0, TR=299.0)
    kw = dict(thickness=1.0, T0=300.0, dt=10.0, maxtime=200.0, nparticles=20000, sigma=1.0)
    sg = kw.pop("sigma")
    g = pkg.rtamc2d(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq), seed=1, sigma_bounce=sg, **kw)
    o = orc.rta_run(m, rect, kind, teq, seed=2, sigma=sg, **kw)
    assert g["Eeff"] == pytest.approx(o["Eeff"], rel=1e-12)
    assert np.isfinite(g["T"]).all()


@pytest.mark.parametrize("what", ["no_particles", "bad_dt", "T0_mismatch"])
def test_rta_rejects_bad_configurations(pkg, ctx, what):
    m = pkg.workloads.material(200, seed=2)
    rect, kind, teq = bar(nbins=3)
    kw = dict(thickness=1.0, T0=300.0, dt=10.0, maxtime=100.0, nparticles=100, seed=1, sigma_bounce=2.0)
    if what == "no_particles":
        kw["nparticles"] = 0
    elif what == "bad_dt":
        kw["dt"] = -10.0
    else:
        kw["T0"] = 310.0
    with pytest.raises(pkg.BteError, match="EINVAL"):
        pkg.RtaRun(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq), **kw)


def test_bert_rejects_a_propagator_of_the_wrong_size(pkg, ctx):
    m, B, P = _bert_inputs(20, seed=2)
    rect, kind, teq, tin = pkg.workloads.ribbon(40.0, 20.0, 4, 300.0, Tinit_first=301.0)
    with pytest.raises(pkg.BteError, match="EINVAL"):
        pkg.BertRun(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq, tin), pkg.Propagator(ctx, P[:19, :19]),
                    thickness=1.0, dt=5.0, maxtime=20.0, Efactor=1000, Tmax=301.0, Tmin=300.0)


def test_spectral_location_out_of_range(pkg, ctx):
    m = pkg.workloads.material(200, seed=2)
    rect, kind, teq = bar(nbins=3)
    r = pkg.RtaRun(ctx, _modes(pkg, ctx, m), pkg.Geometry(ctx, rect, kind, teq), thickness=1.0, T0=300.0, dt=10.0,
                   maxtime=100.0, nparticles=100, seed=1, sigma_bounce=2.0)
    try:
        with pytest.raises(pkg.BteError, match="EINVAL"):
            r.spectral([len(kind)], 10)
    finally:
        r.close()


def test_expm_of_one_by_one_and_zero_matrices(pkg, ctx):
    E, _ = pkg.expm(ctx, np.array([[-2.0]]), 1.5)
    assert E[0, 0] == pytest.approx(np.exp(-3.0), rel=1e-14)
    E, _ = pkg.expm(ctx, np.zeros((5, 5)), 3.0)
    np.testing.assert_allclose(E, np.eye(5), rtol=0, atol=1e-15)
