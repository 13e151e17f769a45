"""Oracle geometry (geometry_2d restated for rectangles) against analytic
slab intersections and hand-built contact layouts."""
import numpy as np
import pytest

import oracle_lib as orc

BOX = np.array([[0.0, 10.0, 0.0, 5.0]])


@pytest.mark.parametrize("x0,y0,vx,vy", [(2.0, 1.0, 1.0, 0.3), (9.0, 4.0, -2.0, -0.1), (5.0, 2.5, 0.0, 1.0),
                                          (5.0, 2.5, 0.2, -3.0), (1.0, 1.0, -0.5, 0.0)])
def test_exit_matches_slab_formula(x0, y0, vx, vy):
    dt = 100.0
    tx = (10 - x0) / vx if vx > 0 else ((0 - x0) / vx if vx < 0 else np.inf)
    ty = (5 - y0) / vy if vy > 0 else ((0 - y0) / vy if vy < 0 else np.inf)
    t = min(tx, ty)
    time, rx, ry, bids = orc.inter_side(BOX, 0, x0, y0, vx, vy, dt)
    assert time == pytest.approx(t, rel=1e-12)
    assert rx == pytest.approx(x0 + t * vx, abs=1e-12) and ry == pytest.approx(y0 + t * vy, abs=1e-12)
    # border order of get_sides for this vertex order: left, bottom, top, right
    side = {0: (vx < 0 and tx <= ty), 1: (vy < 0 and ty <= tx), 2: (vy > 0 and ty <= tx), 3: (vx > 0 and tx <= ty)}
    assert set(bids) == {k for k, v in side.items() if v}


def test_corner_hit_reports_both_sides():
    time, rx, ry, bids = orc.inter_side(BOX, 0, 5.0, 0.0 + 2.5, 5.0, 2.5, 10.0)
    assert time == pytest.approx(1.0) and (rx, ry) == (10.0, 5.0)
    assert sorted(bids) == [2, 3]


def test_particle_on_border_moving_out_exits_at_time_zero():
    """the already_there pass of get_inter_side (geometry_2d.cpp:212-276)"""
    time, rx, ry, bids = orc.inter_side(BOX, 0, 10.0, 2.0, 1.0, 0.5, 3.0)
    assert time == 0.0 and bids == [3]


def test_contacts_of_a_2x2_grid():
    rect = np.array([[0, 1, 0, 1], [1, 2, 0, 1], [0, 1, 1, 2], [1, 2, 1, 2]], float)
    nc, nbr, geo = orc.contacts(rect)
    assert list(nc) == [3, 3, 3, 3]
    # box 0: edge contacts with 1 (normal +x) and 2 (normal +y), point contact with 3
    k = {int(b): i for i, b in enumerate(nbr[0, :3])}
    np.testing.assert_allclose(geo[0, k[1], :2], [1, 0])
    np.testing.assert_allclose(geo[0, k[2], :2], [0, 1])
    np.testing.assert_allclose(geo[0, k[3], :2], np.array([0.5, 0.5]) / np.sqrt(0.5))
    np.testing.assert_allclose(geo[0, k[1], 2:], [1, 0, 1, 1])


def test_unconnected_boxes_raise():
    rect = np.array([[0, 1, 0, 1], [5, 6, 0, 1]], float)
    with pytest.raises(RuntimeError, match="unconnected"):
        orc.contacts(rect)
