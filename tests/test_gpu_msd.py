"""beRTAMC2D on the GPU against a closed form: a 1 K pulse spreading under
RTA scattering in a row of boxes periodic in y (no wall is reached) has the
exact displacement variance of the time-stepped process
(tests/msd_closed_form.py), which in turn reaches the reference's analytic1d
pulse MSD as dt -> 0 (test_oracle_msd.py)."""
import numpy as np
import pytest

from test_oracle_msd import expected_variance, moments_from_props, msd_material, ribbon_row

pytestmark = pytest.mark.gpu


def test_bert_rta_pulse_spreads_as_the_markov_moments(pkg, ctx):
    m = msd_material()
    rect, kind, teq, tin, per, xc = ribbon_row()
    modes = pkg.ModeTable(ctx, m["omega"], m["vx"], m["vy"], m["w"], m["V"], m["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin, periodic=per)
    dt, nsteps, N = 0.5, 60, 1_000_000
    o = pkg.bertamc2d(ctx, modes, geom, None, thickness=1.0, dt=dt, maxtime=dt * (nsteps - 1), max_steps=nsteps,
                      rta=True, Efactor=N, Tmax=301.0, Tmin=300.0, seed=9)
    mean, var, tot = moments_from_props(o["props"], xc)
    ev = expected_variance(m, dt, nsteps, 2.0)
    assert (o["npart"] == o["npart"][0]).all() and o["npart"][0] > 0.9 * N
    np.testing.assert_allclose(tot, tot[0], rtol=1e-9)  # no wall is reached, energy stays
    for n in (5, 10, 20, 40, 60):
        print(f"step {n}: variance {var[n]:.3f} nm^2, closed form {ev[n]:.3f}", flush=True)
        assert abs(var[n] - ev[n]) < 0.01 * ev[n], (n, var[n], ev[n])
    # the centre stays at the pulse box (symmetric spectra carry no drift)
    assert np.abs(mean[1:] - 1.0).max() < 0.05 * np.sqrt(ev[-1])
