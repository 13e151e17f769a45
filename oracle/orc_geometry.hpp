// Oracle (test infrastructure only): restatement of alma::geometry_2d for
// rectangular boxes with the reference's semantics.
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
    double xmin, xmax, ymin, ymax, cx, cy, area;
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
    bool inside(double x, double y) const {
        return x >= xmin && x <= xmax && y >= ymin && y <= ymax;
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
inline void calculate_contacts(std::vector<Box>& sys) {
    for (auto& a : sys) {
        a.contacts.clear();
        for (auto& e : sys) {
            if (e.id == a.id) continue;
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
