"""The oracle's diffuse bounce (diffusive_bouncing, RTA_MC2D.cpp:547-679) and
its corner handling in evolparticle (RTA_MC2D.cpp:1140-1380), pinned by what
the physics and the geometry fix:

* Lambert's cosine law: with one frequency the outgoing mode f has
  probability max(-n.v_f, 0) / sum, n the side's outward normal;
* energy conservation within the Gaussian broadening: modes farther than
  nsigma * hypot(sigma, sigma) in frequency are never reached, and the odds
  of a partner at frequency distance d are exp(-d^2 / (4 sigma^2)) times its
  cosine weight;
* ballistic flights are unaffected by internal box boundaries, through
  edges and exact corners alike (up to the 1e-6 nm nudges);
* a flight into a corner with void on both sides bounces off either side
  with probability 1/2 by symmetry, and into a void corner between two
  boxes it ends on one of them with the mirror-image odds.
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as orc


def ring_modes(n, omega, seed, vmin=0.5, vmax=8.0):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, n)
    sp = rng.uniform(vmin, vmax, n)
    return np.full(n, float(omega)), sp * np.cos(ang), sp * np.sin(ang)


@pytest.mark.parametrize("normal", [(0.0, 1.0), (0.0, -1.0), (1.0, 0.0), (-1.0, 0.0)])
def test_bounce_follows_the_cosine_law(normal):
    om, vx, vy = ring_modes(24, 20.0, seed=5)
    npx, npy = normal
    ns = 200000
    out = orc.lambert_sample(om, vx, vy, 0.25, npx, npy, mode=3, seed=11, ns=ns)
    w = np.maximum(-(npx * vx + npy * vy), 0.0)
    assert (w[out] > 0).all(), "a bounce produced a mode moving into the wall"
    p = w / w.sum()
    cnt = np.bincount(out, minlength=len(om))
    keep = p > 0
    chi = stats.chisquare(cnt[keep], ns * p[keep])
    assert chi.pvalue > 1e-4, chi
    # a uniform choice among the outgoing modes (a dropped cosine) is rejected
    pu = keep / keep.sum()
    assert stats.chisquare(cnt[keep], ns * pu[keep]).pvalue < 1e-12


def test_bounce_never_leaves_the_energy_window():
    a = ring_modes(12, 5.0, seed=1)
    b = ring_modes(12, 9.0, seed=2)
    om, vx, vy = (np.concatenate([x, y]) for x, y in zip(a, b))
    for mode in (0, 17):
        out = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=mode, seed=3, ns=20000)
        assert (om[out] == om[mode]).all()


def test_bounce_odds_across_frequencies_follow_the_broadening():
    sigma, d = 0.25, 0.5
    a = ring_modes(16, 5.0, seed=7)
    b = ring_modes(16, 5.0 + d, seed=8)
    om, vx, vy = (np.concatenate([x, y]) for x, y in zip(a, b))
    w = np.maximum(-vy, 0.0)  # top side, n = (0, 1)
    ratio = np.exp(-d * d / (4 * sigma * sigma))  # normal pdf of d with sd hypot(sigma, sigma)
    p_switch = ratio * w[16:].sum() / (w[:16].sum() + ratio * w[16:].sum())
    ns = 100000
    out = orc.lambert_sample(om, vx, vy, sigma, 0.0, 1.0, mode=2, seed=9, ns=ns)
    got = (out >= 16).mean()
    assert abs(got - p_switch) < 5 * np.sqrt(p_switch * (1 - p_switch) / ns), (got, p_switch)
    # a broadening of sigma instead of hypot(sigma, sigma) would give exp(-2)
    wrong = np.exp(-d * d / (2 * sigma * sigma))
    p_wrong = wrong * w[16:].sum() / (w[:16].sum() + wrong * w[16:].sum())
    assert abs(got - p_wrong) > 20 * np.sqrt(p_wrong * (1 - p_wrong) / ns)


def test_zero_frequency_modes_are_never_reached():
    om, vx, vy = ring_modes(10, 4.0, seed=4)
    om = np.concatenate([om, [0.0]])
    vx = np.concatenate([vx, [0.0]])
    vy = np.concatenate([vy, [-8.0]])  # the most favoured direction, but omega = 0
    out = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=1, seed=2, ns=20000)
    assert (out != 10).all()


def one_mode(vx, vy, extra=()):
    """Mode 0 moves with (vx, vy); extra modes give the bounces partners."""
    om, ex, ey = ring_modes(32, 20.0, seed=12)
    om = np.concatenate([[20.0], om])
    return {"omega": om, "vx": np.concatenate([[vx], ex]), "vy": np.concatenate([[vy], ey])}


GRID4 = np.array([[0, 10, 0, 10], [10, 20, 0, 10], [0, 10, 10, 20], [10, 20, 10, 20]], float)


def follow(mat, rect, kind, box, x, y, dt, seed, max_events=20):
    """Chains single events until the flight ends; returns the visited kinds."""
    kinds = []
    for k in range(max_events):
        r = orc.rta_event(mat, rect, kind, box, x, y, 0, dt, seed=seed + 1000 * k, ns=1)
        kinds.append(orc.EVENT_KINDS[int(r["kind"][0])])
        box, (x, y), dt = int(r["box"][0]), r["xy"][0], float(r["dt"][0])
        if kinds[-1] != "crossing" and kinds[-1] != "nudge":
            return kinds, box, x, y, dt, int(r["mode"][0])
    raise AssertionError(f"flight did not end: {kinds}")


@pytest.mark.parametrize("start,v,dt,end,endbox", [
    ((5.0, 5.0), (2.0, 2.0), 4.0, (13.0, 13.0), 3),   # through the shared corner of four boxes
    ((5.0, 3.0), (2.0, 1.0), 5.0, (15.0, 8.0), 1),    # through an edge
    ((15.0, 15.0), (-2.0, -2.0), 4.0, (7.0, 7.0), 0),  # the corner from the other side
    ((5.0, 5.0), (0.0, 3.0), 3.0, (5.0, 14.0), 2),    # axis-parallel
])
def test_ballistic_flights_ignore_internal_boundaries(start, v, dt, end, endbox):
    mat = one_mode(*v)
    kind = np.zeros(4, np.int32)
    box0 = int(np.flatnonzero((GRID4[:, 0] <= start[0]) & (start[0] <= GRID4[:, 1]) &
                              (GRID4[:, 2] <= start[1]) & (start[1] <= GRID4[:, 3]))[0])
    for seed in range(6):
        kinds, box, x, y, left, mode = follow(mat, GRID4, kind, box0, *start, dt, seed)
        assert kinds[-1] == "inside" and "bounce" not in kinds, kinds
        assert box == endbox and mode == 0
        assert abs(x - end[0]) < 1e-5 and abs(y - end[1]) < 1e-5, (x, y)


def test_closed_corner_bounces_off_either_side_with_equal_odds():
    """A diagonal flight into the corner (10, 10) of a closed square box: the
    two sides are void, each is chosen with probability 1/2 (the approach is
    symmetric under x <-> y), so the outgoing mode follows the mixture of the
    two cosine laws."""
    om, ex, ey = ring_modes(16, 20.0, seed=21)
    # close the table under x <-> y so that both sides see the same partners
    om = np.concatenate([[20.0], om, om])
    vx = np.concatenate([[2.0], ex, ey])
    vy = np.concatenate([[2.0], ey, ex])
    mat = {"omega": om, "vx": vx, "vy": vy}
    rect = np.array([[0, 10, 0, 10]], float)
    ns = 40000
    r = orc.rta_event(mat, rect, np.zeros(1, np.int32), 0, 5.0, 5.0, 0, 4.0, seed=5, ns=ns)
    assert (r["kind"] == 1).all()
    assert np.allclose(r["xy"], 10.0, atol=1e-9)
    wt, wr = np.maximum(-vy, 0.0), np.maximum(-vx, 0.0)
    p = 0.5 * wt / wt.sum() + 0.5 * wr / wr.sum()
    cnt = np.bincount(r["mode"], minlength=len(om))
    keep = p > 0
    assert cnt[~keep].sum() == 0
    assert stats.chisquare(cnt[keep], ns * p[keep]).pvalue > 1e-4
    # always the top side (or always the right side) is rejected
    for q in (wt / wt.sum(), wr / wr.sum()):
        k2 = q > 0
        assert cnt[~k2].sum() > 0 or stats.chisquare(cnt[k2], ns * q[k2]).pvalue < 1e-12


def test_void_corner_between_two_boxes():
    """Boxes 0, 1, 2 of the 2 x 2 grid (no box 3): a diagonal flight from box 0
    into the corner (10, 10) continues in box 1 or box 2 and bounces off the
    void side there (top of box 1, right of box 2), by symmetry each with
    probability 1/2; the particle stays on the geometry and leaves the wall."""
    mat = one_mode(2.0, 2.0)
    rect = GRID4[:3]
    kind = np.zeros(3, np.int32)
    seen = {1: 0, 2: 0}
    R = 400
    for seed in range(R):
        kinds, box, x, y, left, mode = follow(mat, rect, kind, 0, 5.0, 5.0, 4.0, seed)
        assert kinds[-1] == "bounce", kinds
        assert box in (1, 2)
        assert abs(x - 10.0) < 1e-5 and abs(y - 10.0) < 1e-5
        vx, vy = mat["vx"][mode], mat["vy"][mode]
        assert (vy < 0) if box == 1 else (vx < 0)
        seen[box] += 1
    assert abs(seen[1] - R / 2) < 5 * np.sqrt(R / 4), seen


def _three_groups(sig_i, sig_a, sig_b, d=0.5, n=16):
    """mode group i at omega 5 and partner groups a at 5 + d and b at 5 - d,
    each of n modes with the given per-mode broadening"""
    g = [ring_modes(n, w, seed=s) for w, s in ((5.0, 21), (5.0 + d, 22), (5.0 - d, 23))]
    om, vx, vy = (np.concatenate(x) for x in zip(*g))
    sig = np.concatenate([np.full(n, sig_i), np.full(n, sig_a), np.full(n, sig_b)])
    return om, vx, vy, sig


def test_bounce_with_per_mode_sigma_follows_the_pair_width():
    """Per-mode broadening (base_sigma, RTA_MC2D.cpp:566-618): a partner f of
    mode i at frequency distance d has weight N(d; 0, hypot(sigma_i,
    sigma_f)), the normal pdf with its 1/sigma factor, times the cosine
    weight. With a wide group above and a narrow one below, the three
    outcomes follow those odds; dropping the pdf's 1/sigma (which a constant
    sigma would not notice) is rejected."""
    d, si, sa, sb = 0.5, 0.3, 0.6, 0.2
    om, vx, vy, sig = _three_groups(si, sa, sb, d)
    w = np.maximum(-vy, 0.0)   # top side, n = (0, 1)
    pdf = lambda x, s: np.exp(-0.5 * x * x / (s * s)) / (s * np.sqrt(2 * np.pi))
    W = [w[:16].sum(), w[16:32].sum(), w[32:].sum()]
    wt = np.array([W[0] * pdf(0.0, np.hypot(si, si)), W[1] * pdf(d, np.hypot(si, sa)), W[2] * pdf(d, np.hypot(si, sb))])
    ns = 200000
    out = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=2, seed=31, ns=ns, mode_sigma=sig)
    cnt = np.bincount(out // 16, minlength=3)
    assert stats.chisquare(cnt, ns * wt / wt.sum()).pvalue > 1e-4
    # without the 1/sigma of the pdf the odds shift by far more than the noise
    nopre = wt * np.array([np.hypot(si, si), np.hypot(si, sa), np.hypot(si, sb)])
    assert stats.chisquare(cnt, ns * nopre / nopre.sum()).pvalue < 1e-12


def test_bounce_sigma_floor_raises_low_outliers():
    """Per-mode sigma below exp(p25 - 1.5 (p75 - p25)) of log sigma is
    raised to that floor (calc_percentiles_log, isotopic_scattering.cpp:
    31-50, applied at RTA_MC2D.cpp:572-576): a table whose outlier group is
    far below the floor bounces exactly like the same table with that group
    at the floor (same streams); had the floor been skipped, the outliers'
    weights and so the draws would differ."""
    om, vx, vy, sig = _three_groups(0.3, 0.6, 0.2)
    sig = np.concatenate([sig, np.full(8, 1e-4)])          # 8 outliers far below the rest
    extra = ring_modes(8, 5.3, seed=24)
    om, vx, vy = (np.concatenate([a, b]) for a, b in zip((om, vx, vy), extra))
    ls = np.sort(np.log(sig))
    n = len(sig)
    lb = np.exp(ls[25 * n // 100] - 1.5 * (ls[75 * n // 100] - ls[25 * n // 100]))
    assert 1e-4 < lb < 0.2
    floored = np.maximum(sig, lb)
    a = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=2, seed=5, ns=20000, mode_sigma=sig)
    b = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=2, seed=5, ns=20000, mode_sigma=floored)
    assert np.array_equal(a, b)
    # and the outliers do take part in the odds: at their own width the draws differ
    c = orc.lambert_sample(om, vx, vy, 0.25, 0.0, 1.0, mode=2, seed=5, ns=20000,
                           mode_sigma=np.concatenate([sig[:48], np.full(8, 0.2)]))
    assert not np.array_equal(a, c)
