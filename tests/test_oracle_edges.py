"""Boundary cases of the oracle itself: a geometry at equilibrium generates
no particle (beyondRTA_MC2d.cpp:203-270 draws counts from the box
deviations), and a single closed box with adiabatic walls keeps particles
plus pending wall counts constant (every wall hit is re-emitted within the
step, beyondRTA_MC2d.cpp:281-435)."""
import numpy as np

import oracle_lib as orc
from test_oracle_mc import _bert_inputs


def test_oracle_equilibrium_has_no_particles():
    m, B, P = _bert_inputs(30, seed=3)
    rect = np.array([[10.0 * k, 10.0 * (k + 1), 0.0, 30.0] for k in range(4)])
    kind = np.zeros(4, np.int32)
    teq = np.full(4, 300.0)
    o = orc.bert_run(m, rect, kind, teq, teq, thickness=1.0, dt=5.0, maxtime=30.0, P=P, Efactor=10000, Tmax=301.0,
                     Tmin=300.0, seed=1)
    assert o["nsteps"] == 7
    assert (o["npart"] == 0).all() and (o["pending"] == 0).all() and not o["props"].any()


def test_oracle_closed_box_conserves():
    m, B, P = _bert_inputs(30, seed=4)
    rect = np.array([[0.0, 20.0, 0.0, 20.0]])
    o = orc.bert_run(m, rect, np.zeros(1, np.int32), np.array([300.0]), np.array([301.5]), thickness=1.0, dt=5.0,
                     maxtime=50.0, P=P, Efactor=20000, Tmax=301.5, Tmin=300.0, seed=2)
    tot = o["npart"] + o["pending"]
    assert tot[0] > 0 and (tot == tot[0]).all()
