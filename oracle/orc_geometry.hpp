// Oracle (test infrastructure only): restatement of alma::geometry_2d for
// rectangular and convex polygon boxes with the reference's semantics.
//   hull, centre:   geometry_2d.cpp:127-153 (convex hull, area centroid)
//   borders:        geometry_2d.cpp:55-125 (get_sides; sbl = side extended by
//                   1e-6*nd, outward normal np)
//   inside:         geometry_2d.cpp:155-158 (closed region)
//   get_inter_side: geometry_2d.cpp:160-293 (first side with np.v>0 hit by
//                   r0 -> r0+dt*v; corner ties; already_there second pass)
//   contacts:       geometry_2d.cpp:295-381
//   in_corner3:     geometry_2d.cpp:388-406
// Boost.Geometry is replaced by exact segment arithmetic: for axis-aligned
// sides the intersection time is (x_side - x0)/vx, as in the reference's
// max-of-axes formula (geometry_2d.cpp:182-188).
#pragma once
#include "orc_common.hpp"
#include <map>

namespace orc {

struct Border {
    double p1x, p1y, p2x, p2y;  // segment
    double ndx, ndy;            // p2 - p1
    double npx, npy;            // outward unit normal
    double ax, ay, bx, by;      // extended segment (sbl)
    double length() const { return std::hypot(p2x - p1x, p2y - p1y); }
    int Eborder = 0;            // beRTAMC2D wall counter (geometry_2d.hpp:85)
    double Tboxeq = -1;
};

inline double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

// closed-segment intersection of a0->a1 with b0->b1; returns point
inline bool seg_intersect(double a0x, double a0y, double a1x, double a1y,
                          double b0x, double b0y, double b1x, double b1y,
                          double& px, double& py) {
    double dx = a1x - a0x, dy = a1y - a0y, ex = b1x - b0x, ey = b1y - b0y;
    double den = cross2(dx, dy, ex, ey);
    if (den == 0.) return false;  // parallel: the caller excludes np.v == 0
    double wx = b0x - a0x, wy = b0y - a0y;
    double t = cross2(wx, wy, ex, ey) / den;
    double s = cross2(wx, wy, dx, dy) / den;
    if (t < 0. || t > 1. || s < 0. || s > 1.) return false;
    // snap the coordinate that lies on an axis-aligned side exactly
    px = a0x + t * dx;
    py = a0y + t * dy;
    if (ex == 0.) px = b0x;
    if (ey == 0.) py = b0y;
    return true;
}

inline bool point_on_segment(double px, double py, double ax, double ay,
                             double bx, double by) {
    if (cross2(bx - ax, by - ay, px - ax, py - ay) != 0.) return false;
    return px >= std::min(ax, bx) && px <= std::max(ax, bx) &&
           py >= std::min(ay, by) && py <= std::max(ay, by);
}

struct Box {
    int id = 0;
    double xmin, xmax, ymin, ymax, cx, cy, area;  // bounding box for polygons
    std::vector<std::pair<double, double>> hull;  // convex polygons: hull, counter-clockwise
    bool reservoir = false, periodic = false;
    int target = -1;             // periodic: box2translate
    double tx = 0., ty = 0.;     // periodic: translation vector
    double Teq = -1, Treal = -1;
    std::vector<Border> borders;
    std::map<int, Border> contacts;  // contact box id -> shared border

    void make(double x0, double x1, double y0, double y1) {
        xmin = x0; xmax = x1; ymin = y0; ymax = y1;
        cx = 0.5 * (x0 + x1);
        cy = 0.5 * (y0 + y1);
        area = (x1 - x0) * (y1 - y0);
        // vertex order (xmin,ymin),(xmin,ymax),(xmax,ymax),(xmax,ymin); the
        // pair loop of get_sides yields left, bottom, top, right
        double vx[4] = {x0, x0, x1, x1}, vy[4] = {y0, y1, y1, y0};
        int pairs[4][2] = {{0, 1}, {0, 3}, {1, 2}, {2, 3}};
        borders.clear();
        for (auto& pr : pairs) {
            Border b{};
            b.p1x = vx[pr[0]]; b.p1y = vy[pr[0]];
            b.p2x = vx[pr[1]]; b.p2y = vy[pr[1]];
            b.ndx = b.p2x - b.p1x; b.ndy = b.p2y - b.p1y;
            double mx = 0.5 * (b.p1x + b.p2x) - cx, my = 0.5 * (b.p1y + b.p2y) - cy;
            double nx = b.ndy, ny = -b.ndx;
            if (nx * mx + ny * my < 0.) { nx = -nx; ny = -ny; }
            double nn = std::hypot(nx, ny);
            b.npx = nx / nn; b.npy = ny / nn;
            b.ax = b.p1x - 1.0e-6 * b.ndx; b.ay = b.p1y - 1.0e-6 * b.ndy;
            b.bx = b.p2x + 1.0e-6 * b.ndx; b.by = b.p2y + 1.0e-6 * b.ndy;
            borders.push_back(b);
        }
    }
    // geometry_2d(vertices) for a convex polygon (geometry_2d.cpp:127-153):
    // hull, area centroid, area; borders by get_sides' loop over the vertex
    // pairs (i, j >= i) of the given order: a pair is a border when its
    // midpoint moved 1e-4 away from the centre leaves the hull, the normal is
    // the perpendicular pointing away from the centre, a normal already
    // present is skipped (geometry_2d.cpp:55-125)
    void make_poly(const std::vector<double>& vx, const std::vector<double>& vy) {
        std::vector<std::pair<double, double>> p;
        for (size_t i = 0; i < vx.size(); i++) p.emplace_back(vx[i], vy[i]);
        std::sort(p.begin(), p.end());
        p.erase(std::unique(p.begin(), p.end()), p.end());
        hull.clear();
        // monotone chain, lower then upper hull, collinear points dropped
        for (int pass = 0; pass < 2; pass++) {
            const size_t start = hull.size();
            for (size_t q = 0; q < p.size(); q++) {
                const auto& pt = pass ? p[p.size() - 1 - q] : p[q];
                while (hull.size() >= start + 2) {
                    const auto& a = hull[hull.size() - 2];
                    const auto& b = hull[hull.size() - 1];
                    if (cross2(b.first - a.first, b.second - a.second, pt.first - a.first, pt.second - a.second) > 0.)
                        break;
                    hull.pop_back();
                }
                hull.push_back(pt);
            }
            hull.pop_back();
        }
        double A = 0., sx = 0., sy = 0.;
        xmin = ymin = 1e300;
        xmax = ymax = -1e300;
        for (size_t k = 0; k < hull.size(); k++) {
            const auto& a = hull[k];
            const auto& b = hull[(k + 1) % hull.size()];
            const double w = cross2(a.first, a.second, b.first, b.second);
            A += w;
            sx += (a.first + b.first) * w;
            sy += (a.second + b.second) * w;
            xmin = std::min(xmin, a.first); xmax = std::max(xmax, a.first);
            ymin = std::min(ymin, a.second); ymax = std::max(ymax, a.second);
        }
        area = 0.5 * A;
        cx = sx / (3. * A);
        cy = sy / (3. * A);
        borders.clear();
        for (size_t i = 0; i < vx.size(); i++)
            for (size_t j = i; j < vx.size(); j++) {
                if (std::hypot(vx[j] - vx[i], vy[j] - vy[i]) < 1.0e-6) continue;
                const double scx = 0.5 * (vx[i] + vx[j]), scy = 0.5 * (vy[i] + vy[j]);
                double ncx = scx - cx, ncy = scy - cy;
                const double nn = std::hypot(ncx, ncy);
                if (nn == 0.) continue;
                ncx *= 1.0e-4 / nn;
                ncy *= 1.0e-4 / nn;
                if (inside(ncx + scx, ncy + scy)) continue;
                Border b{};
                b.p1x = vx[i]; b.p1y = vy[i]; b.p2x = vx[j]; b.p2y = vy[j];
                b.ndx = b.p2x - b.p1x; b.ndy = b.p2y - b.p1y;
                double nx = b.ndy, ny = -b.ndx;
                if (nx * ncx + ny * ncy < 0.) { nx = -nx; ny = -ny; }
                const double L = std::hypot(nx, ny);
                b.npx = nx / L; b.npy = ny / L;
                b.ax = b.p1x - 1.0e-6 * b.ndx; b.ay = b.p1y - 1.0e-6 * b.ndy;
                b.bx = b.p2x + 1.0e-6 * b.ndx; b.by = b.p2y + 1.0e-6 * b.ndy;
                bool dup = false;
                for (auto& o : borders) dup = dup || std::hypot(o.npx - b.npx, o.npy - b.npy) < 1.0e-6;
                if (!dup) borders.push_back(b);
                if (borders.size() > vx.size()) throw geometry_error("To much borders");
            }
    }
    bool inside(double x, double y) const {
        if (hull.empty()) return x >= xmin && x <= xmax && y >= ymin && y <= ymax;
        for (size_t k = 0; k < hull.size(); k++) {
            const auto& a = hull[k];
            const auto& b = hull[(k + 1) % hull.size()];
            if (cross2(b.first - a.first, b.second - a.second, x - a.first, y - a.second) < 0.) return false;
        }
        return true;
    }
    // get_random_point (geometry_2d.hpp:332-348): uniform in the bounding box
    // (x, then y) until covered by the box
    template <class Rng>
    std::pair<double, double> random_point(Rng& rng) const {
        while (true) {
            const double x = xmin + rng.uniform() * (xmax - xmin);
            const double y = ymin + rng.uniform() * (ymax - ymin);
            if (inside(x, y)) return {x, y};
        }
    }

    struct Hit { double time; double rx, ry; std::vector<int> bids; };

    Hit get_inter_side(double x0, double y0, double vx, double vy, double dt) const {
        double fx = x0 + dt * vx, fy = y0 + dt * vy;
        Hit h{1.0e7, 0, 0, {}};
        bool already_there = true;
        for (int ib = 0; ib < (int)borders.size(); ib++) {
            const Border& b = borders[ib];
            double px, py;
            if (seg_intersect(x0, y0, fx, fy, b.ax, b.ay, b.bx, b.by, px, py) &&
                !point_on_segment(x0, y0, b.ax, b.ay, b.bx, b.by)) {
                if (b.npx * vx + b.npy * vy > 0.) {
                    double t1 = std::fabs(vx) < 1e-6 ? -1 : (px - x0) / vx;
                    double t2 = std::fabs(vy) < 1e-6 ? -1 : (py - y0) / vy;
                    double nt = std::max(t1, t2);
                    if (nt == h.time) h.bids.push_back(ib);
                    if (nt < h.time) {
                        h.bids.clear();
                        h.rx = px; h.ry = py; h.time = nt;
                        already_there = false;
                        h.bids.push_back(ib);
                    }
                }
            }
        }
        if (already_there) {
            double qx = x0 - 1.0e-6 * vx, qy = y0 - 1.0e-6 * vy;
            for (int ib = 0; ib < (int)borders.size(); ib++) {
                const Border& b = borders[ib];
                double px, py;
                if (seg_intersect(qx, qy, fx, fy, b.ax, b.ay, b.bx, b.by, px, py)) {
                    if (b.npx * vx + b.npy * vy > 0.) {
                        double t1 = std::fabs(vx) < 1e-6 ? -1.0e8 : (px - x0) / vx;
                        double t2 = std::fabs(vy) < 1e-6 ? -1.0e8 : (py - y0) / vy;
                        double nt = std::max(t1, t2);
                        if (nt == h.time) h.bids.push_back(ib);
                        if (std::fabs(nt) < h.time) {
                            h.bids.clear();
                            h.rx = px; h.ry = py; h.time = nt;
                            h.bids.push_back(ib);
                        }
                    }
                }
            }
            if (h.time > 1.0e6) throw geometry_error("Error to huge time");
        }
        if (almost_equal(h.time, 0.)) h.time = 0.;
        if (h.time < 0.) throw geometry_error("Bad time");
        return h;
    }
};

// calculate_contacts for rectangles (geometry_2d.cpp:295-381): a shared edge
// of positive length gives a border whose normal points away from this box;
// a single shared point gives a point contact with np = p - centre.
// Shared boundary of two convex polygons with disjoint interiors: the end
// points of the overlaps of collinear side pairs and the meeting points of
// the others, as a segment between the two farthest points or one point;
// two sides crossing inside both is an overlap (an error).
inline bool poly_shared(const Box& a, const Box& e, double tol, double& x0, double& y0, double& x1, double& y1) {
    std::vector<std::pair<double, double>> pts;
    auto add = [&](double x, double y) {
        for (auto& q : pts)
            if (std::hypot(q.first - x, q.second - y) <= tol) return;
        pts.emplace_back(x, y);
    };
    for (auto& ba : a.borders)
        for (auto& be : e.borders) {
            const double dx = ba.ndx, dy = ba.ndy, L = std::hypot(dx, dy);
            const double c1 = cross2(dx, dy, be.p1x - ba.p1x, be.p1y - ba.p1y) / L;
            const double c2 = cross2(dx, dy, be.p2x - ba.p1x, be.p2y - ba.p1y) / L;
            if (std::fabs(c1) <= tol && std::fabs(c2) <= tol) {
                // collinear: overlap of the parameter intervals along ba
                double u1 = ((be.p1x - ba.p1x) * dx + (be.p1y - ba.p1y) * dy) / (L * L);
                double u2 = ((be.p2x - ba.p1x) * dx + (be.p2y - ba.p1y) * dy) / (L * L);
                const double lo = std::max(0., std::min(u1, u2)), hi = std::min(1., std::max(u1, u2));
                if (hi < lo - tol / L) continue;
                add(ba.p1x + lo * dx, ba.p1y + lo * dy);
                add(ba.p1x + hi * dx, ba.p1y + hi * dy);
                continue;
            }
            double px, py;
            if (!seg_intersect(ba.p1x, ba.p1y, ba.p2x, ba.p2y, be.p1x, be.p1y, be.p2x, be.p2y, px, py)) continue;
            const bool end_a = std::hypot(px - ba.p1x, py - ba.p1y) <= tol || std::hypot(px - ba.p2x, py - ba.p2y) <= tol;
            const bool end_e = std::hypot(px - be.p1x, py - be.p1y) <= tol || std::hypot(px - be.p2x, py - be.p2y) <= tol;
            if (!end_a && !end_e) throw geometry_error("overlapping boxes");
            add(px, py);
        }
    if (pts.empty()) return false;
    size_t i0 = 0, i1 = 0;
    double best = -1.;
    for (size_t i = 0; i < pts.size(); i++)
        for (size_t j = i; j < pts.size(); j++) {
            const double d = std::hypot(pts[i].first - pts[j].first, pts[i].second - pts[j].second);
            if (d > best) { best = d; i0 = i; i1 = j; }
        }
    x0 = pts[i0].first; y0 = pts[i0].second; x1 = pts[i1].first; y1 = pts[i1].second;
    return true;
}

inline void calculate_contacts(std::vector<Box>& sys) {
    double scale = 1.;
    for (auto& b : sys) scale = std::max({scale, std::fabs(b.xmin), std::fabs(b.xmax), std::fabs(b.ymin), std::fabs(b.ymax)});
    const double tol = 1.0e-9 * scale;
    for (auto& a : sys) {
        a.contacts.clear();
        for (auto& e : sys) {
            if (e.id == a.id) continue;
            if (!a.hull.empty() || !e.hull.empty()) {
                double x0, y0, x1, y1;
                if (!poly_shared(a, e, tol, x0, y0, x1, y1)) continue;
                Border cb{};
                if (std::hypot(x1 - x0, y1 - y0) < 1.0e-6) {
                    const double px = 0.5 * (x0 + x1), py = 0.5 * (y0 + y1);
                    cb.p1x = cb.p2x = cb.ax = cb.bx = px;
                    cb.p1y = cb.p2y = cb.ay = cb.by = py;
                    const double nx = px - a.cx, ny = py - a.cy, nn = std::hypot(nx, ny);
                    cb.npx = nx / nn; cb.npy = ny / nn;
                    a.contacts[e.id] = cb;
                    continue;
                }
                cb.p1x = x0; cb.p1y = y0; cb.p2x = x1; cb.p2y = y1;
                cb.ndx = x1 - x0; cb.ndy = y1 - y0;
                const double mx = 0.5 * (x0 + x1) - a.cx, my = 0.5 * (y0 + y1) - a.cy;
                double nx = cb.ndy, ny = -cb.ndx;
                if (nx * mx + ny * my < 0.) { nx = -nx; ny = -ny; }
                const double nn = std::hypot(nx, ny);
                cb.npx = nx / nn; cb.npy = ny / nn;
                const double L = std::hypot(cb.ndx, cb.ndy);
                cb.ax = x0 - 1.0e-6 * cb.ndx / L; cb.ay = y0 - 1.0e-6 * cb.ndy / L;
                cb.bx = x1 + 1.0e-6 * cb.ndx / L; cb.by = y1 + 1.0e-6 * cb.ndy / L;
                a.contacts[e.id] = cb;
                continue;
            }
            double x0 = std::max(a.xmin, e.xmin), x1 = std::min(a.xmax, e.xmax);
            double y0 = std::max(a.ymin, e.ymin), y1 = std::min(a.ymax, e.ymax);
            if (x0 > x1 || y0 > y1) continue;  // disjoint
            Border cb{};
            if (x0 == x1 && y0 == y1) {
                cb.p1x = cb.p2x = x0; cb.p1y = cb.p2y = y0;
                cb.ndx = cb.ndy = 0.;
                double nx = x0 - a.cx, ny = y0 - a.cy, nn = std::hypot(nx, ny);
                cb.npx = nx / nn; cb.npy = ny / nn;
                cb.ax = cb.bx = x0; cb.ay = cb.by = y0;
                a.contacts[e.id] = cb;
                continue;
            }
            if (x0 != x1 && y0 != y1) throw geometry_error("overlapping boxes");
            cb.p1x = x0; cb.p1y = y0; cb.p2x = x1; cb.p2y = y1;
            cb.ndx = x1 - x0; cb.ndy = y1 - y0;
            double mx = 0.5 * (x0 + x1) - a.cx, my = 0.5 * (y0 + y1) - a.cy;
            double nx = cb.ndy, ny = -cb.ndx;
            if (nx * mx + ny * my < 0.) { nx = -nx; ny = -ny; }
            double nn = std::hypot(nx, ny);
            cb.npx = nx / nn; cb.npy = ny / nn;
            double L = std::hypot(cb.ndx, cb.ndy);
            cb.ax = x0 - 1.0e-6 * cb.ndx / L; cb.ay = y0 - 1.0e-6 * cb.ndy / L;
            cb.bx = x1 + 1.0e-6 * cb.ndx / L; cb.by = y1 + 1.0e-6 * cb.ndy / L;
            a.contacts[e.id] = cb;
        }
        if (a.contacts.empty() && sys.size() > 1)
            throw geometry_error("There are unconnected boxes");
    }
}

// in_corner3 (geometry_2d.cpp:388-406)
inline bool in_corner3(double x, double y, const std::vector<Box>& sys,
                       std::vector<int>& ids) {
    ids.clear();
    for (auto& g : sys)
        if (g.inside(x, y)) ids.push_back(g.id);
    return ids.size() > 2;
}

// correct_corner_problem (geometry_2d.hpp:524-561)
inline std::pair<int, int> correct_corner_problem(double x, double y, double vx,
                                                  double vy, const std::vector<Box>& sys,
                                                  const std::vector<int>& cand,
                                                  pcg32& rng) {
    std::vector<int> finalist;
    std::map<int, std::vector<int>> bids;
    for (int c : cand) {
        std::vector<int> bb;
        for (int ib = 0; ib < (int)sys[c].borders.size(); ib++) {
            auto& b = sys[c].borders[ib];
            if (point_on_segment(x, y, b.ax, b.ay, b.bx, b.by) &&
                b.npx * vx + b.npy * vy > 0.)
                bb.push_back(ib);
        }
        if (bb.size() <= 1) {
            finalist.push_back(c);
            bids[c] = bb;
        }
    }
    int w = finalist[rng.bounded((uint32_t)finalist.size())];
    if (bids[w].empty()) throw geometry_error("corner without border");
    return {w, bids[w][0]};
}


// geometry_2d::translate (include/geometry_2d.hpp:366-494) on rectangles: a
// particle in periodic box `id` at r0 moves along 1.0005 T to the first
// border of the target box that the translation segment crosses, then into
// the regular neighbour of the target box that holds that point (the shared
// border, the point, or the point nudged 1e-6 along v or back along T).
// Several candidates: those still inside after 1e-6 v, one chosen at random
// (the point is taken from the chosen candidate). Returns the new box.
template <class Rng>
inline int translate(const std::vector<Box>& sys, int id, double& x, double& y, double vx, double vy, Rng& rng) {
    const Box& b = sys[id];
    if (!b.periodic) throw geometry_error("Error: this can only be used in periodic cells");
    const double Tx = 1.0005 * b.tx, Ty = 1.0005 * b.ty;
    const int tib = b.target;
    if (tib < 0 || tib >= (int)sys.size() || almost_equal(std::hypot(Tx, Ty), 0.))
        throw geometry_error("Error: Bad id for mapping of periodic cells");
    const double s0x = x, s0y = y, s1x = x + Tx, s1y = y + Ty;
    double px = 0, py = 0;
    bool found = false;
    for (const Border& br : sys[tib].borders)
        if (seg_intersect(s0x, s0y, s1x, s1y, br.ax, br.ay, br.bx, br.by, px, py)) {
            found = true;
            break;
        }
    if (!found) throw geometry_error("Error in translation, no box found");
    const double tn = std::hypot(b.tx, b.ty), vn = std::hypot(vx, vy);
    const double pTx = px - 1.0e-6 * b.tx / tn, pTy = py - 1.0e-6 * b.ty / tn;
    const double pVx = px + 1.0e-6 * vx / vn, pVy = py + 1.0e-6 * vy / vn;
    std::vector<int> pb;
    std::vector<std::pair<double, double>> rf;
    for (const auto& [ibc, cb] : sys[tib].contacts) {
        double ix, iy;
        const bool here_i = seg_intersect(s0x, s0y, s1x, s1y, cb.ax, cb.ay, cb.bx, cb.by, ix, iy);
        const bool here_T = sys[ibc].inside(px, py);
        const bool here_Tv = sys[ibc].inside(pVx, pVy);
        const bool here_TT = sys[ibc].inside(pTx, pTy);
        if (ibc == tib || !(here_i || here_T || here_Tv || here_TT) || sys[ibc].periodic || sys[ibc].reservoir)
            continue;
        pb.push_back(ibc);
        if (here_i) rf.emplace_back(ix, iy);
        else if (here_T) rf.emplace_back(px, py);
        else if (here_Tv) rf.emplace_back(pVx, pVy);
        else rf.emplace_back(pTx, pTy);
    }
    if (pb.empty()) throw geometry_error("Error in Translation");
    int k = 0;
    if (pb.size() > 1) {
        std::vector<int> keep;
        for (size_t i = 0; i < pb.size(); i++)
            if (sys[pb[i]].inside(rf[i].first + 1.0e-6 * vx, rf[i].second + 1.0e-6 * vy)) keep.push_back((int)i);
        if (keep.empty()) throw geometry_error("Error in Translation");
        k = keep[rng.bounded((uint32_t)keep.size())];
    }
    x = rf[k].first;
    y = rf[k].second;
    return pb[k];
}

}  // namespace orc
