"""The oracle's convex polygon boxes (geometry_2d.cpp:27-153 hull, centre
and get_sides; 160-293 get_inter_side; 295-381 calculate_contacts;
geometry_2d.hpp:332-348 get_random_point) pinned by closed forms, by brute
force on the oracle's own inside test and by the rectangle path: a rectangle
given as a polygon in the reference's vertex order must reproduce the
rectangle's borders, exits and contacts exactly."""
import numpy as np
import pytest

import oracle_lib as orc


def rect_poly(x0, x1, y0, y1):
    # vertex order of the reference's rectangular boxes (beam.xml)
    return np.array([[x0, y0], [x0, y1], [x1, y1], [x1, y0]], float)


def regular(n, r, cx=0.0, cy=0.0, phase=0.3):
    a = phase + 2 * np.pi * np.arange(n) / n
    return np.stack([cx + r * np.cos(a), cy + r * np.sin(a)], 1)


def random_convex(rng, n=9):
    a = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = rng.uniform(5.0, 9.0)
    p = np.stack([r * np.cos(a), r * np.sin(a)], 1) * rng.uniform(0.6, 1.4, 2)
    return p + rng.uniform(-20, 20, 2)


def cross2(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


def fan(p):
    """area and centroid by a triangle fan from the first vertex (the
    vertices in counter-clockwise order)"""
    A, c = 0.0, np.zeros(2)
    for i in range(1, len(p) - 1):
        a = 0.5 * cross2(p[i] - p[0], p[i + 1] - p[0])
        A += a
        c += a * (p[0] + p[i] + p[i + 1]) / 3
    return A, c / A


def test_rectangle_as_polygon_matches_the_rectangle_path():
    rect = np.array([[0.0, 10.0, 0.0, 5.0], [10.0, 25.0, 0.0, 5.0], [0.0, 25.0, 5.0, 9.0]])
    polys = [rect_poly(*r) for r in rect]
    for i in range(3):
        assert np.array_equal(orc.borders(i, rect=rect), orc.borders(i, polys=polys))
    assert np.array_equal(orc.box_info(rect=rect), orc.box_info(polys=polys))
    c1, c2 = orc.contacts(rect), orc.contacts(None, polys=polys)
    for a, b in zip(c1, c2):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(4)
    for _ in range(300):
        i = int(rng.integers(3))
        x0 = rng.uniform(rect[i, 0], rect[i, 1])
        y0 = rng.uniform(rect[i, 2], rect[i, 3])
        ang = rng.uniform(0, 2 * np.pi)
        v = (np.cos(ang), np.sin(ang))
        assert orc.inter_side(rect, i, x0, y0, *v, 100.0) == orc.inter_side(None, i, x0, y0, *v, 100.0, polys=polys)


def test_area_and_centre_closed_forms():
    tri = np.array([[0.0, 0.0], [4.0, 0.0], [0.0, 3.0]])
    hexa = regular(6, 2.0, 7.0, -3.0)
    rng = np.random.default_rng(2)
    rnd = random_convex(rng)
    info = orc.box_info(polys=[tri])[0]
    assert info[0] == pytest.approx(6.0, rel=1e-14) and np.allclose(info[1:3], [4 / 3, 1.0], rtol=1e-14)
    info = orc.box_info(polys=[hexa])[0]
    assert info[0] == pytest.approx(1.5 * np.sqrt(3) * 4.0, rel=1e-13) and np.allclose(info[1:3], [7.0, -3.0], atol=1e-12)
    assert info[7] == 6
    info = orc.box_info(polys=[rnd])[0]
    A, c = fan(rnd)
    assert info[0] == pytest.approx(A, rel=1e-12) and np.allclose(info[1:3], c, rtol=1e-12)
    assert np.allclose(info[3:7], [rnd[:, 0].min(), rnd[:, 0].max(), rnd[:, 1].min(), rnd[:, 1].max()])


def test_borders_are_the_hull_edges_with_outward_normals():
    rng = np.random.default_rng(7)
    for _ in range(5):
        p = random_convex(rng, n=int(rng.integers(3, 12)))
        # shuffled vertex order: the borders are still the hull edges
        q = p[rng.permutation(len(p))]
        b = orc.borders(0, polys=[q])
        assert len(b) == len(p)
        info = orc.box_info(polys=[q])[0]
        for row in b:
            p1, p2, n = row[0:2], row[2:4], row[4:6]
            assert abs(np.hypot(*n) - 1) < 1e-14
            assert np.dot(n, p2 - p1) == pytest.approx(0.0, abs=1e-12 * np.hypot(*(p2 - p1)))
            # every vertex lies behind the border's line
            assert np.all((p - p1) @ n <= 1e-9)
            assert np.dot(0.5 * (p1 + p2) - info[1:3], n) > 0
        # the borders' end points are consecutive hull vertices
        assert sorted(map(tuple, np.round(np.concatenate([b[:, 0:2], b[:, 2:4]]), 12))) == \
            sorted(map(tuple, np.round(np.concatenate([p, p]), 12)))


def test_inside_against_half_planes():
    rng = np.random.default_rng(11)
    p = random_convex(rng)
    pts = np.stack([rng.uniform(p[:, 0].min() - 2, p[:, 0].max() + 2, 4000),
                    rng.uniform(p[:, 1].min() - 2, p[:, 1].max() + 2, 4000)], 1)
    got = orc.inside(0, pts, polys=[p])
    # signed distances to the edges of the counter-clockwise polygon
    d = np.stack([cross2(p[(k + 1) % len(p)] - p[k], pts - p[k]) / np.hypot(*(p[(k + 1) % len(p)] - p[k]))
                  for k in range(len(p))], 1)
    clear = np.abs(d).min(1) > 1e-9
    assert clear.sum() > 3900
    assert np.array_equal(got[clear], (d[clear] >= 0).all(1))
    assert 0.2 < got.mean() < 0.9


def test_exit_time_against_bisection():
    """get_inter_side from an interior point: the exit time equals the sup
    of t with r + t v covered by the box (bisection on the oracle's inside),
    and the returned border holds the exit point"""
    rng = np.random.default_rng(5)
    for trial in range(40):
        p = random_convex(rng, n=int(rng.integers(3, 10)))
        A, c = fan(p)
        w = rng.dirichlet(np.ones(len(p)))
        r0 = 0.5 * c + 0.5 * (w @ p)
        ang = rng.uniform(0, 2 * np.pi)
        v = np.array([np.cos(ang), np.sin(ang)]) * rng.uniform(0.5, 8.0)
        t, rx, ry, bids = orc.inter_side(None, 0, r0[0], r0[1], v[0], v[1], 1e3, polys=[p])
        lo, hi = 0.0, 1e3
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if orc.inside(0, r0 + mid * v, polys=[p])[0]:
                lo = mid
            else:
                hi = mid
        assert t == pytest.approx(lo, rel=1e-9, abs=1e-12), trial
        assert np.allclose([rx, ry], r0 + t * v, atol=1e-9)
        b = orc.borders(0, polys=[p])[bids[0]]
        n = b[4:6]
        assert abs(np.dot([rx, ry] - b[0:2], n)) < 1e-9 and np.dot(n, v) > 0


def test_contacts_of_a_trapezoid_strip_and_a_hexagon_corner():
    # trapezoids side by side sharing slanted edges, and a triangle on top
    xs = [0.0, 4.0, 9.0, 13.0]
    top = [1.0, 5.0, 8.0, 12.0]
    polys = [np.array([[xs[i], 0.0], [xs[i + 1], 0.0], [top[i + 1], 6.0], [top[i], 6.0]]) for i in range(3)]
    polys.append(np.array([[1.0, 6.0], [12.0, 6.0], [6.5, 10.0]]))
    nc, nbr, geo = orc.contacts(None, polys=polys)
    assert list(nc) == [2, 3, 2, 3]
    # box 0 and box 1 share the edge (4,0)-(5,6); the normal of box 0 points
    # to +x side of that edge
    k = list(nbr[0, :nc[0]]).index(1)
    n, p1, p2 = geo[0, k, 0:2], geo[0, k, 2:4], geo[0, k, 4:6]
    assert sorted(map(tuple, np.round([p1, p2], 12))) == [(4.0, 0.0), (5.0, 6.0)]
    assert np.allclose(n, np.array([6.0, -1.0]) / np.hypot(6.0, 1.0))
    # the triangle's contact with box 1 is box 1's whole top edge, normal -y
    k = list(nbr[3, :nc[3]]).index(1)
    assert np.allclose(geo[3, k, 0:2], [0.0, -1.0])
    assert sorted(map(tuple, np.round(geo[3, k, 2:6].reshape(2, 2), 12))) == [(5.0, 6.0), (8.0, 6.0)]
    # three hexagons around one vertex: every pair shares one edge
    r = 3.0
    cs = [np.array([0.0, 0.0]), np.array([1.5 * r, np.sqrt(3) / 2 * r]), np.array([1.5 * r, -np.sqrt(3) / 2 * r])]
    hexes = [regular(6, r, *c, phase=0.0) for c in cs]
    nc, nbr, geo = orc.contacts(None, polys=hexes)
    assert list(nc) == [2, 2, 2]
    for i in range(3):
        for k in range(2):
            j = nbr[i, k]
            L = np.hypot(*(geo[i, k, 4:6] - geo[i, k, 2:4]))
            assert L == pytest.approx(r, rel=1e-9)
            assert np.allclose(geo[i, k, 0:2], (cs[j] - cs[i]) / np.hypot(*(cs[j] - cs[i])), atol=1e-9)


def test_overlapping_polygons_are_an_error():
    a = regular(5, 3.0)
    b = regular(5, 3.0, 2.0, 0.5)
    with pytest.raises(RuntimeError, match="overlapping"):
        orc.contacts(None, polys=[a, b])


def test_random_points_are_uniform_in_the_polygon():
    p = np.array([[0.0, 0.0], [10.0, 0.0], [3.0, 6.0]])
    xy = orc.random_points(0, 40000, 9, polys=[p])
    assert orc.inside(0, xy, polys=[p]).all()
    A, c = fan(p)
    # the mean of a uniform point is the centroid; per-axis standard error
    se = xy.std(0) / np.sqrt(len(xy))
    assert np.all(np.abs(xy.mean(0) - c) < 3.5 * se)
    # a fraction of the area: the sub-triangle below y = 3 holds 3/4 of it
    assert abs((xy[:, 1] < 3.0).mean() - 0.75) < 3.5 * np.sqrt(0.75 * 0.25 / len(xy))
