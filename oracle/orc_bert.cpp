// Oracle (test infrastructure only): restatement of beRTAMC2D
// (src/beyondRTA_MC2d.cpp) for rectangular geometries and one material.
//   calc_Eeff                  beyondRTA_MC2d.cpp:100-128, main 2044-2070
//   init_particles (outTref)   beyondRTA_MC2d.cpp:211-269
//   adiabatic_diffuse_walls    beyondRTA_MC2d.cpp:281-435
//   Isothermal_walls           beyondRTA_MC2d.cpp:446-585
//   evolparticle               beyondRTA_MC2d.cpp:708-970
//   ScatteringRTA              beyondRTA_MC2d.cpp:983-1062
//   Scattering                 beyondRTA_MC2d.cpp:1073-1210
//   propagator_H               collision_operator.hpp:241-323, .cpp:854-916
//   lirp, propagator selection collision_operator.cpp:918-929,
//                              beyondRTA_MC2d.cpp:1131-1166
//   calculate_gradientT        geometry_2d.cpp:524-576
//   source_term_gradient       beyondRTA_MC2d.cpp:597-676, advected 2169-2180
//   delete_particles           beyondRTA_MC2d.cpp:1214-1220
//   histogram / cancellation   beyondRTA_MC2d.cpp:2277-2370
//   calculate_prop (histogram) beyondRTA_MC2d.cpp:1535-1602
//   step loop                  beyondRTA_MC2d.cpp:2138-2425
// Deviations applied identically on the GPU path: per-particle PCG32 streams
// instead of default-seeded per-thread pcg64 (beyondRTA_MC2d.cpp:1992-1993);
// an explicit dead flag instead of the mode (0,0) sentinel; gradient
// generators are visited in box order (the reference iterates an
// unordered_map); a lower_bound past the last cumulative entry picks the last
// mode (the reference reads one past the end there).
#include "orc_geometry.hpp"
#include <omp.h>
#include <map>
#include <array>
#include <limits>

namespace orc {

// propagator_H (collision_operator.cpp:854-870, hpp:256-262, cpp:893-904)
struct PropH {
    std::vector<double> signed_cumm;
    double abssum = 0;
    PropH(std::vector<double> sc, double abs_) : signed_cumm(std::move(sc)), abssum(abs_) {}
    explicit PropH(const double* b, int n) : signed_cumm(n) {
        abssum = 0;
        for (int i = 0; i < n; i++) abssum += std::fabs(b[i]);
        double acc = 0;
        for (int i = 0; i < n; i++) {
            acc += std::fabs(b[i]);
            signed_cumm[i] = b[i] < 0. ? -acc / abssum : acc / abssum;
        }
    }
    // returns (sign, index) with index relative to the first column mode
    std::pair<int, int> get(double R) const {
        size_t im = std::lower_bound(signed_cumm.begin(), signed_cumm.end(), R,
                                     [](double a, double b) { return std::fabs(a) < std::fabs(b); }) -
                    signed_cumm.begin();
        if (im >= signed_cumm.size()) im = signed_cumm.size() - 1;
        return {signed_cumm[im] > 0. ? 1 : -1, (int)im};
    }
};

// lirp (collision_operator.cpp:918-929): linear interpolation of the signed
// cumulatives and of the column sums
static PropH lirp(const PropH& H1, const PropH& H2, double T1, double T2, double Tx) {
    std::vector<double> sc(H1.signed_cumm.size());
    for (size_t i = 0; i < sc.size(); i++)
        sc[i] = H1.signed_cumm[i] + (Tx - T1) * (H2.signed_cumm[i] - H1.signed_cumm[i]) / (T2 - T1);
    return PropH(std::move(sc), H1.abssum + (Tx - T1) * (H2.abssum - H1.abssum) / (T2 - T1));
}

// The pair of tabulated temperatures used for a box temperature Tx that is
// not tabulated. rule 0 is Scattering's choice (beyondRTA_MC2d.cpp:1139-1164):
// T1 = the first key not below Tx (map lower_bound), T2 = the key after it,
// allowed when T2 exists and Tx >= the lowest key; the line through (T1, T2)
// is then evaluated at Tx < T1, on the signed cumulatives (lirp). rule 1
// brackets Tx (T1 < Tx < T2 adjacent) and interpolates the matrices.
static std::pair<double, double> prop_pair(const std::vector<double>& keys, double Tx, int rule) {
    auto it = std::lower_bound(keys.begin(), keys.end(), Tx);
    if (rule == 0) {
        if (it != keys.end() && it + 1 != keys.end() && Tx >= keys.front()) return {*it, *(it + 1)};
    } else if (it != keys.begin() && it != keys.end()) {
        return {*(it - 1), *it};
    }
    throw value_error("Extrapolation is not allowed: limit Tref between the min and max temperatures "
                      "used in Propagator calculation");
}

// (1 - f) P1 + f P2 entry by entry
static std::vector<double> interp_matrix(const std::vector<double>& P1, const std::vector<double>& P2, double f) {
    std::vector<double> P(P1.size());
    for (size_t i = 0; i < P.size(); i++) P[i] = P1[i] + f * (P2[i] - P1[i]);
    return P;
}

struct BertSetup {
    Modes m;
    std::vector<Box> sys;
    double thickness = 1.0, dt = 1.0, maxtime = 10.0;
    int ngamma = 0;              // leading modes not in P (3 in the reference)
    std::vector<double> P;       // (n-ngamma)^2 column-major, used at every Teq
    // <propagator T file> children (beyondRTA_MC2d.cpp:1355-1376): when not
    // empty, replaces P; ascending temperatures
    std::vector<double> PT;
    std::vector<std::vector<double>> Ps;
    int prop_rule = 0;           // prop_pair rule
    // <gradient x y> (beyondRTA_MC2d.cpp:1328-1333, 1441-1460): the same
    // gradient [K/nm] in every regular box instead of calculate_gradientT
    bool fixed_gradient = false;
    double gx = 0., gy = 0.;
    bool rta = false;
    double Ed = -1.;             // fixed particle energy, or
    uint64_t Efactor = 0;        // number of particles for calc_Eeff
    double Tmax = 0, Tmin = 0;
    int pcancel_each = 4;
    // <init_distribution box file> (beyondRTA_MC2d.cpp:1411-1417): a given
    // signed distribution per regular box (n - ngamma entries)
    std::map<int, std::vector<double>> init_dist;

    double Eeff = 0;
};

static double calc_Eeff(const Modes& m, double Th, double Tc, double Tr, double vol, uint64_t np) {
    long double dh = 0, dc = 0;
    for (int i = 0; i < m.n; i++) {
        if (almost_equal(0., m.omega[i])) continue;
        dh += 1e12 * hbar * m.omega[i] * std::fabs(bose_einstein(m.omega[i], Th) - bose_einstein(m.omega[i], Tr));
        dc += 1e12 * hbar * m.omega[i] * std::fabs(bose_einstein(m.omega[i], Tc) - bose_einstein(m.omega[i], Tr));
    }
    return (double)(vol * (dc + dh) / m.nq / np);
}

struct BPart {
    double x, y;
    int mode, box, sign;
    bool dead;
};

// calculate_gradientT (geometry_2d.cpp:524-576): for every regular box the
// minimum-norm least-squares solution g of h_i . g = Teq_i - Teq_0 over its
// regular contact boxes i (h_i: centre differences). The 2-column system is
// solved through its singular value decomposition (one rotation of the
// columns); singular values below 2 eps sigma_max count as zero.
std::vector<std::array<double, 2>> gradient_T(const std::vector<Box>& sys) {
    std::vector<std::array<double, 2>> g(sys.size(), {0., 0.});
    for (auto& s : sys) {
        if (s.reservoir || s.periodic) continue;
        std::vector<double> hx, hy, df;
        for (auto& [c, b] : s.contacts) {
            if (sys[c].reservoir || sys[c].periodic) continue;
            df.push_back(sys[c].Teq - s.Teq);
            hx.push_back(sys[c].cx - s.cx);
            hy.push_back(sys[c].cy - s.cy);
        }
        const size_t m = df.size();
        if (m == 0) continue;
        double a = 0, b = 0, c = 0;  // A^T A = [[a, c], [c, b]]
        for (size_t i = 0; i < m; i++) {
            a += hx[i] * hx[i];
            b += hy[i] * hy[i];
            c += hx[i] * hy[i];
        }
        // right singular vectors v1 = (cs, sn), v2 = (-sn, cs): the
        // eigenvectors of A^T A, at angle atan2(2c, a - b) / 2
        const double th = 0.5 * std::atan2(2. * c, a - b);
        const double cs = std::cos(th), sn = std::sin(th);
        // rotated columns u1 = A v1, u2 = A v2
        double s1 = 0, s2 = 0, p1 = 0, p2 = 0;
        for (size_t i = 0; i < m; i++) {
            const double u1 = cs * hx[i] + sn * hy[i];
            const double u2 = -sn * hx[i] + cs * hy[i];
            s1 += u1 * u1;
            s2 += u2 * u2;
            p1 += u1 * df[i];
            p2 += u2 * df[i];
        }
        const double smax = std::sqrt(std::max(s1, s2));
        const double tol = 2. * std::numeric_limits<double>::epsilon() * (double)std::min<size_t>(m, 2) * smax;
        double gxv = 0, gyv = 0;
        if (std::sqrt(s1) > tol) { gxv += cs * p1 / s1; gyv += sn * p1 / s1; }
        if (std::sqrt(s2) > tol) { gxv += -sn * p2 / s2; gyv += cs * p2 / s2; }
        g[s.id] = {gxv, gyv};
    }
    return g;
}

// Nabla_T_distribution (sampling.cpp:118-164) for an in-plane gradient at T:
// pmf |(-g.v) kernel(omega, T)|, sign of -g.v (>= 0 is +), cumulsum kB sum
static GridDist nabla_T_signed(const Modes& m, double gx, double gy, double T) {
    GridDist d;
    d.nq = m.nq;
    std::vector<double> pmf(m.n, 0.);
    d.signs.assign(m.n, 1);
    double acc = 0;
    for (int i = 0; i < m.n; i++) {
        if (m.omega[i] == 0.) continue;
        const double vgt = -(gx * m.vx[i] + gy * m.vy[i]);
        pmf[i] = std::fabs(vgt * bose_einstein_kernel(m.omega[i], T));
        d.signs[i] = vgt >= 0. ? 1 : -1;
        acc += pmf[i];
    }
    d.cumulsum = kB * acc;
    d.fill_cumulative(pmf);
    return d;
}

// evolparticle (beyondRTA_MC2d.cpp:708-970)
static void bert_evol(BertSetup& S, BPart& p, double vx, double vy, double dt,
                      pcg32& rng, std::vector<std::vector<int>>& Eb, int depth = 0) {
    auto& sys = S.sys;
    if (depth > 100000) throw geometry_error("runaway recursion");
    if (sys[p.box].reservoir) { p.dead = true; return; }
    // beyondRTA_MC2d.cpp:729-737: translate and continue the flight
    if (sys[p.box].periodic) p.box = translate(sys, p.box, p.x, p.y, vx, vy, rng);
    Box& box = sys[p.box];
    double nx = p.x + dt * vx, ny = p.y + dt * vy;
    if (box.inside(nx, ny)) { p.x = nx; p.y = ny; return; }
    Box::Hit hit;
    try {
        hit = box.get_inter_side(p.x, p.y, vx, vy, dt);
    } catch (const geometry_error&) {
        double dx = box.cx - p.x, dy = box.cy - p.y, dn = std::hypot(dx, dy);
        p.x += 1.0e-6 * dx / dn;
        p.y += 1.0e-6 * dy / dn;
        bert_evol(S, p, vx, vy, dt, rng, Eb, depth + 1);
        return;
    }
    if (hit.time < 0) throw geometry_error("Error in kinematics");
    double left = dt - hit.time;
    if (almost_equal(left, 0.)) left = 0.;
    double rx = hit.rx, ry = hit.ry;
    double vn = std::hypot(vx, vy);
    double ex = rx + 1.0e-6 * vx / vn, ey = ry + 1.0e-6 * vy / vn;
    std::vector<int> pboxes;
    for (auto& [c, bc] : box.contacts) {
        if (c == p.box) continue;
        double d = bc.npx * vx + bc.npy * vy;
        if ((sys[c].inside(rx, ry) || sys[c].inside(ex, ey)) && !almost_equal(d, 0.) && d > 0.)
            pboxes.push_back(c);
    }
    if (pboxes.empty()) {
        std::vector<int> cids;
        if (in_corner3(rx, ry, sys, cids)) {
            auto cs = correct_corner_problem(rx, ry, vx, vy, sys, cids, rng);
            p.box = cs.first;
            p.x = rx; p.y = ry;
            Eb[p.box][cs.second] += p.sign;
        } else {
            int ib;
            if (hit.bids.size() > 1) {
                std::vector<int> voidb;
                for (int tvb : hit.bids) {
                    auto& mb = box.borders[tvb];
                    double cx = rx + 1.0e-4 * mb.npx, cy = ry + 1.0e-4 * mb.npy;
                    bool isvoid = true;
                    for (auto& [c, bc] : box.contacts)
                        if (sys[c].inside(cx, cy)) { isvoid = false; break; }
                    if (isvoid) voidb.push_back(tvb);
                }
                if (voidb.empty()) throw geometry_error("Error in void borders");
                ib = voidb[rng.bounded((uint32_t)voidb.size())];
            } else {
                ib = hit.bids[0];
            }
            Eb[p.box][ib] += p.sign;
        }
        p.dead = true;
        return;
    }
    if (pboxes.size() == 1) {
        p.box = pboxes[0];
        p.x = rx; p.y = ry;
        bert_evol(S, p, vx, vy, left, rng, Eb, depth + 1);
        return;
    }
    std::vector<int> pbox2;
    double tau = 1.0e-6;
    while (true) {
        double cx = rx + tau * vx, cy = ry + tau * vy;
        for (int pb : pboxes)
            if (sys[pb].inside(cx, cy)) pbox2.push_back(pb);
        if (!pbox2.empty()) break;
        tau /= 10.0;
        if (tau < 1.0e-15) {
            double dx = box.cx - rx, dy = box.cy - ry, dn = std::hypot(dx, dy);
            p.x = rx + 1.0e-6 * dx / dn;
            p.y = ry + 1.0e-6 * dy / dn;
            bert_evol(S, p, vx, vy, left, rng, Eb, depth + 1);
            return;
        }
    }
    p.box = pbox2[rng.bounded((uint32_t)pbox2.size())];
    p.x = rx; p.y = ry;
    bert_evol(S, p, vx, vy, left, rng, Eb, depth + 1);
}

// Runs the full beRTAMC2D loop. props: [nsteps+1][nbox][3] (ed, jx, jy) with
// row 0 the initial state (beyondRTA_MC2d.cpp:2105-2117); npart: [nsteps+1].
// Returns the number of steps done (<= max_steps).
int bert_run(BertSetup& S, uint64_t seed, int nthreads, int max_steps,
             double* props, int64_t* npart, int32_t* H_last, double* Eeff_out, int64_t* pending) {
    auto& sys = S.sys;
    const Modes& m = S.m;
    int nb = (int)sys.size();
    int ncol = m.n - S.ngamma;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double vuc = m.V * S.thickness / m.lattz;
    std::vector<double> vols(nb), volsf(nb);
    for (int i = 0; i < nb; i++) {
        vols[i] = sys[i].area * S.thickness;
        volsf[i] = vols[i] / vuc;
    }
    // Eeff (beyondRTA_MC2d.cpp:2026-2070)
    if (S.Ed < 0.) {
        double maxT = 0, minT = 1.416784e+32, meanT = 0;
        int nb_ = 0;
        for (auto& s : sys) {
            if (s.periodic) continue;
            meanT += s.Teq;
            nb_++;
            minT = std::min(minT, s.Teq);
            maxT = std::max(maxT, s.Teq);
        }
        if (S.Tmax != 0) { maxT = S.Tmax; minT = S.Tmin; }
        meanT /= nb_;
        double e = 1e300;
        for (int i = 0; i < nb; i++) {
            if (sys[i].periodic || sys[i].reservoir) continue;
            e = std::min(e, calc_Eeff(m, maxT, minT, meanT, volsf[i], S.Efactor));
        }
        S.Eeff = e;
    } else {
        S.Eeff = S.Ed;
    }
    const double Eeff = S.Eeff;
    if (Eeff_out) *Eeff_out = Eeff;
    // propagator_H per column for every regular box temperature: the
    // tabulated one, or lirp of the pair prop_pair picks. The reference makes
    // this choice when a particle scatters; here every regular box is
    // checked once before the first step.
    std::map<double, std::vector<PropH>> hist_by_T;
    std::vector<const std::vector<PropH>*> box_hist(nb, nullptr);
    if (!S.rta) {
        auto columns = [&](const std::vector<double>& P) {
            std::vector<PropH> h;
            for (int c = 0; c < ncol; c++) h.emplace_back(&P[(size_t)c * ncol], ncol);
            return h;
        };
        if (S.PT.empty()) {
            hist_by_T[0.] = columns(S.P);
            for (int b = 0; b < nb; b++) box_hist[b] = &hist_by_T[0.];
        } else {
            std::map<double, std::vector<PropH>> tab;
            for (size_t t = 0; t < S.PT.size(); t++) tab[S.PT[t]] = columns(S.Ps[t]);
            for (auto& s : sys) {
                if (s.reservoir || s.periodic) continue;
                if (!hist_by_T.count(s.Teq)) {
                    if (tab.count(s.Teq)) {
                        hist_by_T[s.Teq] = tab[s.Teq];
                    } else {
                        const auto pr = prop_pair(S.PT, s.Teq, S.prop_rule);
                        std::vector<PropH> h;
                        if (S.prop_rule == 1) {
                            const auto& P1 = S.Ps[std::find(S.PT.begin(), S.PT.end(), pr.first) - S.PT.begin()];
                            const auto& P2 = S.Ps[std::find(S.PT.begin(), S.PT.end(), pr.second) - S.PT.begin()];
                            h = columns(interp_matrix(P1, P2, (s.Teq - pr.first) / (pr.second - pr.first)));
                        } else {
                            const auto& H1 = tab[pr.first];
                            const auto& H2 = tab[pr.second];
                            for (int c = 0; c < ncol; c++) h.push_back(lirp(H1[c], H2[c], pr.first, pr.second, s.Teq));
                        }
                        hist_by_T[s.Teq] = std::move(h);
                    }
                }
                box_hist[s.id] = &hist_by_T[s.Teq];
            }
        }
    }
    // gradient generators (source_term_gradient): boxes whose gradient is
    // not almost zero, each with its Nabla_T distribution at the box Teq
    std::vector<int> gbox;
    std::vector<GridDist> ggen;
    {
        std::vector<std::array<double, 2>> grad;
        if (S.fixed_gradient) {
            grad.assign(nb, {0., 0.});
            for (auto& s : sys)
                if (!s.reservoir && !s.periodic) grad[s.id] = {S.gx, S.gy};
        } else {
            grad = gradient_T(sys);
        }
        for (int b = 0; b < nb; b++) {
            if (almost_equal(std::hypot(grad[b][0], grad[b][1]), 0.)) continue;
            gbox.push_back(b);
            ggen.push_back(nabla_T_signed(m, grad[b][0], grad[b][1], sys[b].Teq));
        }
    }
    std::vector<GridDist> rta_sampler;
    std::map<double, GridDist> rta_by_T;
    if (S.rta)
        for (auto& s : sys)
            if (!rta_by_T.count(s.Teq)) rta_by_T[s.Teq] = be_derivative(m, s.Teq);

    pcg32 hostrng(seed, 0xb0a7u);  // the serial generators' stream
    std::vector<BPart> parts;
    // init_particles from given distributions (beyondRTA_MC2d.cpp:142-193,
    // ref_distribution sampling.cpp:70-114): Ntogenerate(Vbox/Vucell, Eeff)
    // particles at uniform points, mode and sign from the distribution. The
    // reference's second init_particles (from Tinit, below) clears the list
    // first (beyondRTA_MC2d.cpp:2096-2105, 218), so given distributions never
    // survive there; both sets are kept here.
    for (auto& [ib, fd] : S.init_dist) {
        if (ib < 0 || ib >= (int)sys.size()) throw input_error("init_distribution box out of range");
        const Box& s = sys[ib];
        if (s.reservoir || s.periodic) continue;
        GridDist d = ref_distribution(m, fd, S.ngamma);
        size_t ngen = n_to_generate((s.area * S.thickness / vuc) * d.cumulsum / (Eeff * m.nq), hostrng);
        for (size_t i = 0; i < ngen; i++) {
            const auto [x, y] = s.random_point(hostrng);
            size_t pos = d.sample(hostrng);
            parts.push_back({x, y, (int)pos, s.id, d.signs[pos], false});
        }
    }
    // init_particles from Tinit != Teq (beyondRTA_MC2d.cpp:211-269)
    for (auto& s : sys) {
        if (s.reservoir || s.periodic) continue;
        if (almost_equal(s.Teq, s.Treal)) continue;
        GridDist d = out_tref(m, s.Treal, s.Teq);
        size_t ngen = n_to_generate((s.area * S.thickness / vuc) * d.cumulsum / (Eeff * m.nq), hostrng);
        for (size_t i = 0; i < ngen; i++) {
            const auto [x, y] = s.random_point(hostrng);
            size_t pos = d.sample(hostrng);
            parts.push_back({x, y, (int)pos, s.id, d.signs[pos], false});
        }
    }
    auto calc_prop = [&](const std::vector<std::vector<int>>& H, int row) {
        for (int b = 0; b < nb; b++) {
            double ed = 0, fx = 0, fy = 0;
            if (!(sys[b].reservoir || sys[b].periodic)) {
                long long s = 0;
                for (int k = 0; k < m.n; k++) {
                    s += H[b][k];
                    fx += H[b][k] * m.vx[k];
                    fy += H[b][k] * m.vy[k];
                }
                ed = Eeff / vols[b] * 1.0e27 * (double)s;
                fx *= Eeff / vols[b] * 1.0e30;
                fy *= Eeff / vols[b] * 1.0e30;
            }
            props[((size_t)row * nb + b) * 3 + 0] = ed;
            props[((size_t)row * nb + b) * 3 + 1] = fx;
            props[((size_t)row * nb + b) * 3 + 2] = fy;
        }
    };
    // row 0: properties of the initial particles (calculate_prop over the
    // particle list, beyondRTA_MC2d.cpp:1494-1532)
    {
        std::vector<std::vector<int>> H0(nb, std::vector<int>(m.n, 0));
        for (auto& p : parts) H0[p.box][p.mode] += p.sign;
        calc_prop(H0, 0);
        npart[0] = (int64_t)parts.size();
        if (pending) pending[0] = 0;
    }
    // adiabatic wall spline tables J(T) per (box, border)
    std::vector<double> Ts(100);
    for (int i = 0; i < 100; i++) Ts[i] = 10.0 + 10.0 * i;
    std::map<std::pair<int, int>, CubicSpline1D> interp;
    std::vector<std::vector<int>> Eb(nb, std::vector<int>(4, 0));
    double time = 0.;
    int istep = 0;
    uint64_t stepcount = 0;
    while (time <= S.maxtime && istep < max_steps) {
        // 1. advection
        {
            int nth = omp_get_max_threads();
            std::vector<std::vector<std::vector<int>>> Ebt(nth, std::vector<std::vector<int>>(nb, std::vector<int>(4, 0)));
            std::string err;
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)parts.size(); i++) {
                int tid = omp_get_thread_num();
                pcg32 rng(seed ^ 0x9e3779b97f4a7c15ULL, (stepcount << 32) ^ (uint64_t)i);
                double vx, vy;
                get_v(m, parts[i].mode, vx, vy);
                try {
                    bert_evol(S, parts[i], vx, vy, S.dt, rng, Ebt[tid]);
                } catch (const std::exception& e) {
#pragma omp critical
                    err = e.what();
                }
            }
            if (!err.empty()) throw geometry_error(err);
            for (auto& e : Ebt)
                for (int b = 0; b < nb; b++)
                    for (int k = 0; k < 4; k++) Eb[b][k] += e[b][k];
        }
        auto advect_new = [&](std::vector<BPart>& np_, uint64_t salt) {
            for (size_t i = 0; i < np_.size(); i++) {
                pcg32 rng(seed ^ salt, (stepcount << 32) ^ (uint64_t)i);
                double vx, vy;
                get_v(m, np_[i].mode, vx, vy);
                double rdt = S.dt * rng.uniform();
                bert_evol(S, np_[i], vx, vy, rdt, rng, Eb);
            }
            parts.insert(parts.end(), np_.begin(), np_.end());
        };
        // 2. source_term_gradient: Ntogenerate(dt, vol, Eeff) particles per
        // generator box at uniform points of the box, mode and sign from its
        // Nabla_T distribution, advected U*dt
        {
            std::vector<BPart> np_;
            for (size_t k = 0; k < gbox.size(); k++) {
                const Box& s = sys[gbox[k]];
                const GridDist& g = ggen[k];
                size_t ng = n_to_generate(S.dt * volsf[s.id] * g.cumulsum / (Eeff * m.nq), hostrng);
                for (size_t i = 0; i < ng; i++) {
                    const auto [x, y] = s.random_point(hostrng);
                    size_t pos = g.sample(hostrng);
                    np_.push_back({x, y, (int)pos, s.id, g.signs[pos], false});
                }
            }
            advect_new(np_, 0x9ad5ULL);
        }
        // 3. Isothermal walls (reservoir injection)
        {
            std::vector<BPart> np_;
            for (auto& s : sys) {
                if (!s.reservoir) continue;
                for (auto& [c, b] : s.contacts) {
                    if (sys[c].reservoir || sys[c].periodic) continue;
                    if (almost_equal(b.length(), 0.)) continue;
                    if (almost_equal(s.Teq, sys[c].Teq)) continue;
                    GridDist g = isothermal_wall(m, s.Teq, sys[c].Teq, b.npx, b.npy);
                    double spf = b.length() * S.thickness / vuc;
                    size_t ng = n_to_generate(S.dt * spf * g.cumulsum / (Eeff * m.nq), hostrng);
                    for (size_t k = 0; k < ng; k++) {
                        double u = hostrng.uniform();
                        double x = b.p1x + u * b.ndx, y = b.p1y + u * b.ndy;
                        if (!sys[c].inside(x, y)) {
                            double dx = sys[c].cx - x, dy = sys[c].cy - y, dn = std::hypot(dx, dy);
                            x += 1.0e-6 * dx / dn;
                            y += 1.0e-6 * dy / dn;
                        }
                        size_t pos = g.sample(hostrng);
                        np_.push_back({x, y, (int)pos, c, g.signs[pos], false});
                    }
                }
            }
            advect_new(np_, 0x15a0ULL);
        }
        // 4. adiabatic diffuse walls
        {
            std::vector<BPart> np_;
            for (auto& s : sys) {
                if (s.reservoir || s.periodic) continue;
                for (int ib = 0; ib < (int)s.borders.size(); ib++) {
                    auto& b = s.borders[ib];
                    auto key = std::make_pair(s.id, ib);
                    if (!interp.count(key)) {
                        std::vector<double> J(100);
                        for (int t = 0; t < 100; t++) {
                            GridDist g = isothermal_wall(m, Ts[t], s.Teq, -b.npx, -b.npy);
                            J[t] = wall_flux(g, vuc);
                        }
                        interp[key] = CubicSpline1D(J, Ts);
                    }
                    if (Eb[s.id][ib] != 0) {
                        double Jb = Eeff * Eb[s.id][ib] / S.dt / (b.length() * S.thickness);
                        double Tb = interp[key].interpolate(Jb);
                        GridDist g = isothermal_wall(m, Tb, s.Teq, -b.npx, -b.npy);
                        size_t ng = (size_t)(std::abs(Eb[s.id][ib]) + 1.0e-6);
                        for (size_t k = 0; k < ng; k++) {
                            double u = hostrng.uniform();
                            double x = b.p1x + u * b.ndx, y = b.p1y + u * b.ndy;
                            size_t pos = g.sample(hostrng);
                            np_.push_back({x, y, (int)pos, s.id, g.signs[pos], false});
                        }
                        Eb[s.id][ib] = 0;
                    }
                }
            }
            advect_new(np_, 0xad1aULL);
        }
        // 5. delete
        parts.erase(std::remove_if(parts.begin(), parts.end(), [](const BPart& p) { return p.dead; }), parts.end());
        // 6. scattering
        if (S.rta) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)parts.size(); i++) {
                pcg32 rng(seed ^ 0x5ca7ULL, (stepcount << 32) ^ (uint64_t)i);
                auto& p = parts[i];
                double R = rng.uniform();
                if (R < 1 - std::exp(-m.w[p.mode] * S.dt))
                    p.mode = (int)rta_by_T.at(sys[p.box].Teq).sample(rng);
            }
        } else {
            std::vector<BPart> out;
            out.reserve(parts.size() * 2);
            int nth = omp_get_max_threads();
            std::vector<std::vector<BPart>> outs(nth);
            bool badmode = false;
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)parts.size(); i++) {
                int tid = omp_get_thread_num();
                pcg32 rng(seed ^ 0x5ca7ULL, (stepcount << 32) ^ (uint64_t)i);
                long pending = 1;
                const BPart parent = parts[i];
                if (parent.mode - S.ngamma < 0) {
                    badmode = true;
                    continue;
                }
                while (pending > 0) {
                    pending--;
                    int col = parent.mode - S.ngamma;
                    auto sol = (*box_hist[parent.box])[col].get(rng.uniform());
                    BPart c = parent;
                    c.mode = sol.second + S.ngamma;
                    int ns = (parent.sign * sol.first) < 0 ? -1 : 1;
                    c.sign = ns;
                    outs[tid].push_back(c);
                    if (ns * parent.sign < 0) pending += 2;
                }
            }
            if (badmode) throw value_error("This particle should be deleted");
            for (auto& o : outs) out.insert(out.end(), o.begin(), o.end());
            parts.swap(out);
        }
        // 7. histogram
        std::vector<std::vector<int>> H(nb, std::vector<int>(m.n, 0));
        for (auto& p : parts) H[p.box][p.mode] += p.sign;
        // 8. cancellation (sorted by (box, mode, sign); keep the first |h|
        // particles of sign(h) per (box, mode)). The reference sorts with
        // tbb::parallel_sort, which is not stable: the order within one
        // (box, mode, sign) group, and so which particles survive, is
        // unspecified. The oracle takes a uniformly random subset (a shuffle
        // before a stable sort).
        if (istep % S.pcancel_each == 0) {
            pcg32 shuf(seed ^ 0xca9ce1ULL, stepcount);
            for (size_t i = parts.size(); i > 1; i--) {
                const uint64_t r = ((uint64_t)shuf.next() << 32) | shuf.next();
                const size_t j = (size_t)(r % i);
                std::swap(parts[i - 1], parts[j]);
            }
            std::stable_sort(parts.begin(), parts.end(), [](const BPart& a, const BPart& b) {
                if (a.box != b.box) return a.box < b.box;
                if (a.mode != b.mode) return a.mode < b.mode;
                return a.sign < b.sign;
            });
            auto Hc = H;
            std::vector<BPart> kept;
            kept.reserve(parts.size());
            for (auto& p : parts) {
                if (sys[p.box].reservoir || sys[p.box].periodic) continue;
                int& h = Hc[p.box][p.mode];
                if (h == 0) continue;
                if ((p.sign < 0) != (h < 0)) continue;
                h -= p.sign;
                kept.push_back(p);
            }
            parts.swap(kept);
        }
        istep++;
        stepcount++;
        time += S.dt;
        calc_prop(H, istep);
        npart[istep] = (int64_t)parts.size();
        if (pending) {
            long long pe = 0;
            for (int b = 0; b < nb; b++)
                for (int k = 0; k < 4; k++) pe += Eb[b][k];
            pending[istep] = pe;
        }
        if (H_last)
            for (int b = 0; b < nb; b++)
                for (int k = 0; k < m.n; k++) H_last[(size_t)b * m.n + k] = H[b][k];
    }
    return istep;
}

}  // namespace orc
