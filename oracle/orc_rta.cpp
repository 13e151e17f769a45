// Oracle (test infrastructure only): restatement of the RTAMC2D transient
// branch (src/RTA_MC2D.cpp) for rectangular geometries and one material.
//   Init_generator          RTA_MC2D.cpp:272-533 (isothermal-wall part)
//   diffusive_bouncing      RTA_MC2D.cpp:547-679
//   process_segment         RTA_MC2D.cpp:699-769, get_bin_index 140-148
//   spectral decomposition  RTA_MC2D.cpp:61-63 (gamma_pdf), 718-766,
//                           1930-1945, 1978-2052, 2349-2405
//   evolparticle            RTA_MC2D.cpp:990-1386 (DMM and periodic parts out)
//   history loop            RTA_MC2D.cpp:2228-2317
//   post-processing         RTA_MC2D.cpp:2326-2435
// Deviations, each forced by the test setting and applied identically to
// the GPU path: one PCG32 stream per particle (the reference copies one
// pcg64 into every thread, RTA_MC2D.cpp:2131-2134); u in (0,1] for random_dt
// (sampling.hpp:44-45 allows u = 0); the random void-border choice uses
// [0,size) (RTA_MC2D.cpp:1234-1235 is inclusive); the dead-particle sentinel
// is an explicit flag instead of mode (0,0); the Lambert broadening uses a
// constant sigma per mode because a synthetic table has no q-grid.
// Three extensions have no reference counterpart (the reference's void sides
// are diffuse only, RTA_MC2D.cpp:1203-1260); they follow their definitions:
// a scattering-rate factor wscale on every w; at side s a specular
// reflection with probability spec[s] (the mode goes to its mirror partner,
// v_x negated on the left and right sides, v_y on the bottom and top),
// diffuse otherwise; and the score-function derivative of every tally with
// respect to spec[s]: each particle carries sum over its wall events on s of
// d/dp log P(outcome) (1/p specular, -1/(1-p) diffuse), and every segment
// adds contr * score[s] to a per-side tally.
#include "orc_geometry.hpp"
#include <limits>
#include <omp.h>

namespace orc {

using LambertTable = std::vector<std::vector<std::pair<int, double>>>;

// The broadening floor of the reference: modes below exp(p25 - 1.5 (p75 -
// p25)) of log sigma are raised to it (calc_percentiles_log,
// isotopic_scattering.cpp:31-50, applied at RTA_MC2D.cpp:572-576).
static std::vector<double> floored_sigma(std::vector<double> sig) {
    const size_t n = sig.size();
    std::vector<double> ls(n);
    for (size_t i = 0; i < n; i++) ls[i] = sig[i] == 0. ? std::numeric_limits<double>::lowest() : std::log(sig[i]);
    std::sort(ls.begin(), ls.end());
    const double p0 = ls[25 * n / 100], p1 = ls[75 * n / 100];
    const double lb = std::exp(p0 - 1.5 * (p1 - p0));
    for (auto& v : sig) v = std::max(v, lb);
    return sig;
}

// diffusive_bouncing table (RTA_MC2D.cpp:556-626): for mode i every mode f
// with nonzero omega within nsigma * hypot(sigma_i, sigma_f) in frequency,
// weighted by the normal pdf of |omega_i - omega_f| with that width. The
// per-mode sigma_i is base_sigma (RTA_MC2D.cpp:566-570) after the floor of
// floored_sigma; a table without it (a synthetic one has no q-grid) uses
// sigma for every mode.
static LambertTable lambert_table(const Modes& m, double sigma, const std::vector<double>& mode_sigma) {
    LambertTable tab(m.n);
    std::vector<double> bs = mode_sigma.empty() ? std::vector<double>(m.n, sigma) : floored_sigma(mode_sigma);
    for (int i = 0; i < m.n; i++) {
        if (almost_equal(0., m.omega[i])) continue;
        for (int f = 0; f < m.n; f++) {
            if (almost_equal(0., m.omega[f])) continue;
            const double sig = std::hypot(bs[i], bs[f]);
            double d = std::fabs(m.omega[i] - m.omega[f]);
            if (d <= nsigma * sig)
                tab[i].push_back({f, std::exp(-0.5 * d * d / (sig * sig)) / (sig * std::sqrt(2 * PI))});
        }
    }
    return tab;
}

// diffusive_bouncing::bounce (RTA_MC2D.cpp:635-678): f with probability
// proportional to max(-n.v_f, 0) * table(i, f), n the side's outward normal
static int lambert_bounce(const Modes& m, const LambertTable& L, double npx, double npy, int mode, pcg32& rng) {
    const auto& tm = L[mode];
    double cum = 0;
    std::vector<std::pair<int, double>> states;
    states.reserve(tm.size());
    for (auto& [f, econ] : tm) {
        double vx = m.vx[f], vy = m.vy[f];
        if (almost_equal(vx, 0.)) vx = 0.;
        if (almost_equal(vy, 0.)) vy = 0.;
        double vn = -(npx * vx + npy * vy);
        if (vn < 0. || almost_equal(vn, 0.)) continue;
        cum += vn * econ;
        states.push_back({f, vn * econ});
    }
    if (states.empty()) throw geometry_error("no diffuse partner");
    double s = rng.uniform(), c2 = 0;
    for (auto& [f, tr] : states) {
        c2 += tr / cum;
        if (c2 >= s) return f;
    }
    return states.back().first;
}

struct RtaSetup {
    Modes m;
    std::vector<Box> sys;
    double thickness = 1.0, T0 = 300.0, dt = 1.0, maxtime = 10.0;
    bool ballistic = false;
    double sigma = 0.25;  // Lambert broadening per mode [rad/ps]
    std::vector<double> mode_sigma;  // per-mode base_sigma (empty: sigma for all)
    double wscale = 1.0;  // scattering-rate factor
    double spec[4] = {0., 0., 0., 0.};  // specular probability: left, bottom, top, right
    int gmask = 0;        // sides whose specularity is differentiated (0 < p < 1)
    std::vector<int> mirror;  // [2][n]: v_x negated, v_y negated; -1 none
    std::vector<double> grad_out;  // dT/dp [side][tbin][box] of the last run
    // <spectral> (RTA_MC2D.cpp:1930-1945, 1978-2052): boxes with a spectral
    // decomposition, nomega ticks, per-mode broadening (base_sigma; empty:
    // sigma for every mode)
    std::vector<int> spec_boxes;
    int nomega = 0;
    std::vector<double> spec_sigma;
    std::vector<int> spec_loc;        // per box: location index or -1
    std::vector<double> omegagrid;    // tick centres
    std::vector<double> spec_out[4];  // results: dT, jx, jy [loc][tbin][omega]; fd [loc][tbin][q]
    std::vector<double> bsig;         // broadening per mode after the outlier floor

    // derived
    int ntime = 0;
    std::vector<double> timegrid;
    GridDist sampler;
    double cv = 0;
    struct Gen { int idr, idi; Border b; GridDist d; double weight; };
    std::vector<Gen> gens;
    std::vector<double> gen_cum;
    double einfo1 = 0, Eeff = 0;
    // Lambert table: candidates within nsigma*sqrt(2)*sigma, with pdf values
    std::vector<std::vector<std::pair<int, double>>> lambert;

    void prepare_spectral() {
        spec_loc.assign(sys.size(), -1);
        if (spec_boxes.empty()) return;
        if (nomega < 1) throw input_error("spectral resolution must be at least one tick");
        for (size_t l = 0; l < spec_boxes.size(); l++) {
            const int b = spec_boxes[l];
            if (b < 0 || b >= (int)sys.size()) throw input_error("spectral location out of range");
            spec_loc[b] = (int)l;
        }
        // omega grid: LinSpaced(nomega, 0, maxomega) + delta/2, delta =
        // maxomega / nomega (RTA_MC2D.cpp:2040-2045)
        double maxomega = 0;
        for (int i = 0; i < m.n; i++) maxomega = std::max(maxomega, m.omega[i]);
        const double delta = maxomega / nomega;
        omegagrid.resize(nomega);
        for (int i = 0; i < nomega; i++)
            omegagrid[i] = (nomega > 1 ? maxomega * i / (nomega - 1) : maxomega) + 0.5 * delta;
        bsig.resize(m.n);
        for (int i = 0; i < m.n; i++)
            bsig[i] = !spec_sigma.empty() ? spec_sigma[i] : (!mode_sigma.empty() ? mode_sigma[i] : sigma);
        // outlier floor exp(p25 - 1.5 (p75 - p25)) of log sigma
        // (calc_percentiles_log, isotopic_scattering.cpp:31-50)
        std::vector<double> ls(m.n);
        for (int i = 0; i < m.n; i++)
            ls[i] = bsig[i] == 0. ? std::numeric_limits<double>::lowest() : std::log(bsig[i]);
        std::sort(ls.begin(), ls.end());
        const double p0 = ls[25 * (size_t)m.n / 100], p1 = ls[75 * (size_t)m.n / 100];
        const double lb = std::exp(p0 - 1.5 * (p1 - p0));
        for (auto& v : bsig) v = std::max(v, lb);
    }

    void prepare(uint64_t nparticles) {
        prepare_spectral();
        ntime = (int)std::floor(maxtime / dt + 1.0e-6) + 1;
        maxtime = (ntime - 1) * dt;
        timegrid.resize(ntime);
        for (int i = 0; i < ntime; i++)
            timegrid[i] = ntime > 1 ? (double)i * maxtime / (ntime - 1) : 0.;
        sampler = be_derivative(m, T0);
        cv = calc_cv(m, T0);
        for (auto& s : sys) {
            if (s.reservoir || s.periodic) continue;
            if (s.Teq != T0)
                throw input_error("Tref in xml is different from provided T0");
            if (s.Treal != s.Teq)
                throw input_error("Initial temperature out of reference is not allowed");
        }
        // isothermal generators, ordered like the std::unordered_map is
        // irrelevant: selection is by energy weight
        double vuc = m.V * thickness / m.lattz;
        gens.clear();
        for (auto& s : sys) {
            if (!s.reservoir) continue;
            for (auto& [c, b] : s.contacts) {
                if (sys[c].reservoir || sys[c].periodic) continue;
                if (almost_equal(b.length(), 0.)) continue;
                if (almost_equal(s.Teq, sys[c].Teq)) continue;
                Gen g{s.id, c, b, isothermal_wall(m, s.Teq, sys[c].Teq, b.npx, b.npy), 0.};
                double jd = wall_flux(g.d, vuc);
                g.weight = std::fabs(jd) * b.length() * thickness;
                gens.push_back(g);
            }
        }
        double tot = 0;
        for (auto& g : gens) tot += g.weight;
        einfo1 = tot;
        gen_cum.clear();
        double cum = 0;
        for (auto& g : gens) {
            cum += g.weight / tot;
            gen_cum.push_back(cum);
        }
        Eeff = einfo1 / (double)nparticles;
        // Lambert table (RTA_MC2D.cpp:556-626) with sigma_i = sigma for all modes
        lambert = lambert_table(m, sigma, mode_sigma);
        // mirror partners: same omega, one velocity component negated, exact match
        mirror.assign(2 * (size_t)m.n, -1);
        bool anyspec = false;
        for (int k = 0; k < 4; k++) {
            if (!(spec[k] >= 0. && spec[k] <= 1.)) throw input_error("specularity must lie in [0, 1]");
            anyspec = anyspec || spec[k] > 0.;
            if (!(spec[k] > 0. && spec[k] < 1.)) gmask &= ~(1 << k);
        }
        if (anyspec) {
            std::map<std::tuple<double, double, double>, int> idx;
            for (int i = 0; i < m.n; i++) idx.emplace(std::make_tuple(m.omega[i], m.vx[i], m.vy[i]), i);
            for (int i = 0; i < m.n; i++) {
                auto a = idx.find(std::make_tuple(m.omega[i], -m.vx[i], m.vy[i]));
                auto b = idx.find(std::make_tuple(m.omega[i], m.vx[i], -m.vy[i]));
                if (a != idx.end()) mirror[i] = a->second;
                if (b != idx.end()) mirror[(size_t)m.n + i] = b->second;
            }
        }
    }

    // side index of a rectangle side from its outward normal
    static int side_of(double npx, double npy) {
        return npx < 0. ? 0 : (npy < 0. ? 1 : (npy > 0. ? 2 : 3));
    }

    // a void side: specular with probability spec[s], else the diffuse bounce;
    // the score of a differentiated side gets d/dp log P(outcome)
    int reflect(double npx, double npy, int mode, pcg32& rng, double* score) const {
        const int s = side_of(npx, npy);
        const double p = spec[s];
        bool specular = p >= 1.;
        if (p > 0. && p < 1.) {
            specular = rng.uniform() < p;
            if (score && ((gmask >> s) & 1)) score[s] += specular ? 1. / p : -1. / (1. - p);
        }
        if (!specular) return bounce(npx, npy, mode, rng);
        const int f = mirror[(s == 0 || s == 3 ? 0 : (size_t)m.n) + mode];
        if (f < 0) throw input_error("specular reflection of a mode without a mirror partner");
        return f;
    }

    int bounce(double npx, double npy, int mode, pcg32& rng) const {
        return lambert_bounce(m, lambert, npx, npy, mode, rng);
    }
};

struct Part {
    double x, y, t;
    int mode, box, sign;
    bool dead;
    double score[4] = {0., 0., 0., 0.};
};

struct Tally {
    int nbox, nt;
    std::vector<double> gz, jx, jy;  // [box][tbin]
    std::vector<double> dgz;         // [side][box][tbin], score-weighted gz
    // spectral: [loc][tbin][omega] and [loc][tbin][q]
    std::vector<double> gzw, jxw, jyw, fdq;
    Tally(int nb, int nt_) : nbox(nb), nt(nt_), gz(nb * nt_, 0.), jx(nb * nt_, 0.), jy(nb * nt_, 0.) {}
    void spectral(size_t nloc, int nomega, int nq) {
        gzw.assign(nloc * nt * nomega, 0.);
        jxw.assign(nloc * nt * nomega, 0.);
        jyw.assign(nloc * nt * nomega, 0.);
        fdq.assign(nloc * nt * nq, 0.);
    }
};

// gamma_pdf (RTA_MC2D.cpp:61-63): density of the Gamma distribution with
// shape k and scale theta at x (boost gamma_p_derivative(k, x/theta)/theta)
inline double gamma_pdf(double k, double theta, double x) {
    if (x <= 0.) return 0.;
    const double z = x / theta;
    return std::exp((k - 1.) * std::log(z) - z - std::lgamma(k)) / theta;
}

static inline int bin_index(const std::vector<double>& grid, double target) {
    long idx = (long)(std::lower_bound(grid.begin(), grid.end(), target) - grid.begin()) - 1;
    long hi = (long)grid.size() - 2;
    return (int)std::clamp(idx, 0L, hi);
}

// process_segment, transient overload (RTA_MC2D.cpp:699-769), with the
// Gamma-broadened spectral tallies of a spectral location (718-766)
static void process_segment(const RtaSetup& S, Tally& T, int box, double t0,
                            double t1, int sign, double vx, double vy, int mode,
                            const double* score = nullptr) {
    int b0 = bin_index(S.timegrid, t0), b1 = bin_index(S.timegrid, t1);
    double delta = t1 - t0;
    const int loc = S.spec_loc.empty() ? -1 : S.spec_loc[box];
    std::vector<double> gw;
    if (loc >= 0) {
        const double om = S.m.omega[mode], sg = S.bsig[mode];
        const double k = om * om / sg, theta = sg / om;
        gw.resize(S.nomega);
        for (int i = 0; i < S.nomega; i++) gw[i] = gamma_pdf(k, theta, S.omegagrid[i]);
    }
    for (int i = b0; i <= b1; i++) {
        double lo = std::max(S.timegrid[i] - t0, 0.);
        double hi = std::min(S.timegrid[i + 1] - t0, delta);
        double contr = sign * std::max(hi - lo, 0.);
        T.gz[box * T.nt + i] += contr;
        T.jx[box * T.nt + i] += vx * contr;
        T.jy[box * T.nt + i] += vy * contr;
        if (S.gmask && score)
            for (int q = 0; q < 4; q++)
                if ((S.gmask >> q) & 1) T.dgz[((size_t)q * T.nbox + box) * T.nt + i] += contr * score[q];
        if (loc >= 0) {
            const size_t o = ((size_t)loc * T.nt + i) * S.nomega;
            for (int w = 0; w < S.nomega; w++) {
                T.gzw[o + w] += contr * gw[w];
                T.jxw[o + w] += vx * contr * gw[w];
                T.jyw[o + w] += vy * contr * gw[w];
            }
            const int nb = S.m.n / S.m.nq;
            T.fdq[((size_t)loc * T.nt + i) * S.m.nq + mode / nb] += contr / S.m.omega[mode];
        }
    }
}

struct Counters { uint64_t events = 0, crossings = 0, nudges = 0; };

// Init_generator (RTA_MC2D.cpp:436-484): wall chosen by energy weight, a
// uniform point on its contact border, mode and sign from the wall pmf.
static Part init_particle(const RtaSetup& S, pcg32& rng) {
    double R = rng.uniform();
    size_t gi = std::lower_bound(S.gen_cum.begin(), S.gen_cum.end(), R) - S.gen_cum.begin();
    if (gi >= S.gens.size()) gi = S.gens.size() - 1;
    const auto& g = S.gens[gi];
    double u = rng.uniform();
    double x = g.b.p1x + u * g.b.ndx, y = g.b.p1y + u * g.b.ndy;
    const Box& bi = S.sys[g.idi];
    if (!bi.inside(x, y)) {
        double dx = bi.cx - x, dy = bi.cy - y, dn = std::hypot(dx, dy);
        x += 1.0e-6 * dx / dn;
        y += 1.0e-6 * dy / dn;
        if (!bi.inside(x, y)) throw geometry_error("Generated particle out of insertion box");
    }
    size_t pos = g.d.sample(rng);
    return Part{x, y, 0., (int)pos, g.idi, g.d.signs[pos], false};
}

// evolparticle, transient (RTA_MC2D.cpp:990-1386), single material
static void evolparticle(const RtaSetup& S, Part& p, double vx, double vy,
                         double& dt, pcg32& rng, Tally& T, Counters& C) {
    const auto& sys = S.sys;
    const Box& box = sys[p.box];
    if (box.periodic) {
        // RTA_MC2D.cpp:1034-1043: translate and return with dt unchanged
        p.box = translate(sys, p.box, p.x, p.y, vx, vy, rng);
        return;
    }
    if (box.reservoir) {
        p.dead = true;
        dt = 0.;
        C.events++;
        return;
    }
    double nx = p.x + dt * vx, ny = p.y + dt * vy;
    if (box.inside(nx, ny)) {
        p.x = nx;
        p.y = ny;
        process_segment(S, T, p.box, p.t, p.t + dt, p.sign, vx, vy, p.mode, p.score);
        p.t += dt;
        if (!S.ballistic) p.mode = (int)S.sampler.sample(rng);
        C.events++;
        dt = 0.;
        return;
    }
    Box::Hit hit;
    try {
        hit = box.get_inter_side(p.x, p.y, vx, vy, dt);
    } catch (const geometry_error&) {
        double dx = box.cx - p.x, dy = box.cy - p.y, dn = std::hypot(dx, dy);
        p.x += 1.0e-6 * dx / dn;
        p.y += 1.0e-6 * dy / dn;
        C.nudges++;
        return;
    }
    double tf = hit.time;
    double left = dt - tf;
    if (almost_equal(left, 0.)) left = 0.;
    dt = left;
    double rx = hit.rx, ry = hit.ry;
    double vn = std::hypot(vx, vy);
    double ex = rx + 1.0e-6 * vx / vn, ey = ry + 1.0e-6 * vy / vn;
    std::vector<int> pboxes;
    for (auto& [c, bc] : box.contacts) {
        if (c == p.box) continue;
        double d = bc.npx * vx + bc.npy * vy;
        bool parallel = almost_equal(d, 0.);
        bool oriented = d > 0.;
        if ((sys[c].inside(rx, ry) || sys[c].inside(ex, ey)) && !parallel && oriented)
            pboxes.push_back(c);
    }
    if (pboxes.empty()) {
        process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy, p.mode, p.score);
        p.t += tf;
        int bid = p.box;
        std::vector<int> cids;
        std::vector<int> bids = hit.bids;
        if (in_corner3(rx, ry, sys, cids)) {
            auto cs = correct_corner_problem(rx, ry, vx, vy, sys, cids, rng);
            bid = cs.first;
            double dx = sys[bid].cx - rx, dy = sys[bid].cy - ry, dn = std::hypot(dx, dy);
            auto h2 = sys[bid].get_inter_side(rx + 1.0e-6 * dx / dn, ry + 1.0e-6 * dy / dn, vx, vy, 1.0e18);
            rx = h2.rx;
            ry = h2.ry;
            bids = h2.bids;
            p.box = bid;
        }
        int iborder;
        if (bids.size() > 1) {
            std::vector<int> voidb;
            for (int tvb : bids) {
                auto& mb = sys[p.box].borders[tvb];
                double cx = rx + 1.0e-4 * mb.npx, cy = ry + 1.0e-4 * mb.npy;
                bool isvoid = true;
                for (auto& [c, bc] : sys[p.box].contacts)
                    if (sys[c].inside(cx, cy)) { isvoid = false; break; }
                if (isvoid) voidb.push_back(tvb);
            }
            if (voidb.empty()) throw geometry_error("Error in void borders");
            iborder = voidb[rng.bounded((uint32_t)voidb.size())];
        } else {
            iborder = bids[0];
        }
        auto& ub = sys[p.box].borders[iborder];
        p.mode = S.reflect(ub.npx, ub.npy, p.mode, rng, p.score);
        p.x = rx;
        p.y = ry;
        C.events++;
        dt = 0.;
        return;
    }
    if (pboxes.size() == 1) {
        p.x = rx;
        p.y = ry;
        process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy, p.mode, p.score);
        p.t += tf;
        p.box = pboxes[0];
        C.crossings++;
        return;
    }
    process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy, p.mode, p.score);
    p.t += tf;
    p.x = rx;
    p.y = ry;
    std::vector<int> pbox2;
    double cx = rx + 1.0e-6 * vx, cy = ry + 1.0e-6 * vy;
    for (int pb : pboxes)
        if (sys[pb].inside(cx, cy)) pbox2.push_back(pb);
    int newid;
    if (pbox2.empty()) {
        newid = pboxes[rng.bounded((uint32_t)pboxes.size())];
        double dx = sys[newid].cx - rx, dy = sys[newid].cy - ry, dn = std::hypot(dx, dy);
        if (!sys[newid].periodic) {
            p.x += 1.0e-6 * dx / dn;
            p.y += 1.0e-6 * dy / dn;
        }
    } else {
        newid = pbox2[rng.bounded((uint32_t)pbox2.size())];
    }
    p.box = newid;
    C.crossings++;
}

// History loop (RTA_MC2D.cpp:2228-2317) + post-processing (2326-2435).
// Outputs are [time][box] like the CSV rows (without the time column).
int rta_run(RtaSetup& S, uint64_t nparticles, uint64_t seed, int nthreads,
            double* T_out, double* jx_out, double* jy_out, uint64_t* counters,
            double* raw_gz) {
    S.prepare(nparticles);
    int nb = (int)S.sys.size(), nt = S.ntime - 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int nth = omp_get_max_threads();
    std::vector<Tally> tallies(nth, Tally(nb, nt));
    const size_t nloc = S.spec_boxes.size();
    if (nloc)
        for (auto& t : tallies) t.spectral(nloc, S.nomega, S.m.nq);
    if (S.gmask)
        for (auto& t : tallies) t.dgz.assign((size_t)4 * nb * nt, 0.);
    std::vector<Counters> cnt(nth);
    const double maxtime = S.maxtime;
    std::string err;
#pragma omp parallel for schedule(dynamic, 64)
    for (long long i = 0; i < (long long)nparticles; i++) {
        int tid = omp_get_thread_num();
        try {
        pcg32 rng(seed, (uint64_t)i);
        Part p = init_particle(S, rng);
        double time = 0.;
        while (time < maxtime) {
            double vx, vy;
            get_v(S.m, p.mode, vx, vy);
            double w = S.m.w[p.mode] * S.wscale;
            if (w == 0.) throw value_error("invalid scattering rate");
            double u = 1.0 - rng.uniform();
            double dt_ = -std::log(u) / w;
            if (time + dt_ > maxtime) dt_ = maxtime - time + 1.0e-8;
            while (!almost_equal(dt_, 0.))
                evolparticle(S, p, vx, vy, dt_, rng, tallies[tid], cnt[tid]);
            if (p.dead) break;
            time = p.t;
        }
        } catch (const std::exception& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw geometry_error(err);
    // dT/dp [side][tbin][box]: the score-weighted tallies post-processed like T
    S.grad_out.assign((size_t)4 * nt * nb, 0.);
    if (S.gmask) {
        const double vcorr = S.thickness / S.m.lattz;
        for (int q = 0; q < 4; q++)
            for (int b = 0; b < nb; b++) {
                const double vbox = S.sys[b].area * S.thickness;
                double acc = 0.;
                for (int k = 0; k < nt; k++) {
                    double v = 0.;
                    for (int t = 0; t < nth; t++) v += tallies[t].dgz[((size_t)q * nb + b) * nt + k];
                    acc += v * S.Eeff / vbox;
                    S.grad_out[((size_t)q * nt + k) * nb + b] = acc / (S.cv / vcorr);
                }
            }
    }
    std::vector<double> gz(nb * nt, 0.), jx(nb * nt, 0.), jy(nb * nt, 0.);
    Counters tot;
    for (int t = 0; t < nth; t++) {
        for (int k = 0; k < nb * nt; k++) {
            gz[k] += tallies[t].gz[k];
            jx[k] += tallies[t].jx[k];
            jy[k] += tallies[t].jy[k];
        }
        tot.events += cnt[t].events;
        tot.crossings += cnt[t].crossings;
        tot.nudges += cnt[t].nudges;
    }
    if (raw_gz)
        for (int k = 0; k < nb * nt; k++) raw_gz[k] = gz[k];
    // spectral post-processing (RTA_MC2D.cpp:2349-2405): scale, cumulative
    // sum over time, deltaT = gz_omega / (Cv / vcorr)
    if (nloc) {
        const int nw = S.nomega, nq = S.m.nq;
        Tally& A = tallies[0];
        for (int t = 1; t < nth; t++)
            for (size_t k = 0; k < A.gzw.size(); k++) {
                A.gzw[k] += tallies[t].gzw[k];
                A.jxw[k] += tallies[t].jxw[k];
                A.jyw[k] += tallies[t].jyw[k];
            }
        for (int t = 1; t < nth; t++)
            for (size_t k = 0; k < A.fdq.size(); k++) A.fdq[k] += tallies[t].fdq[k];
        const double vcorr = S.thickness / S.m.lattz;
        for (size_t l = 0; l < nloc; l++) {
            const double vbox = S.sys[S.spec_boxes[l]].area * S.thickness;
            for (int i = 0; i < nt; i++) {
                for (int w = 0; w < nw; w++) {
                    const size_t o = (l * nt + i) * nw + w, p = (l * nt + (i > 0 ? i - 1 : 0)) * nw + w;
                    A.gzw[o] *= S.Eeff / vbox;
                    A.jxw[o] *= S.Eeff * 1e30 / vbox;
                    A.jyw[o] *= S.Eeff * 1e30 / vbox;
                    if (i > 0) {
                        A.gzw[o] += A.gzw[p];
                        A.jxw[o] += A.jxw[p];
                        A.jyw[o] += A.jyw[p];
                    }
                }
                for (int q = 0; q < nq; q++) {
                    const size_t o = (l * nt + i) * nq + q;
                    A.fdq[o] *= S.Eeff * vcorr * nq / (vbox * 1.0e12 * hbar);
                    if (i > 0) A.fdq[o] += A.fdq[o - nq];
                }
            }
            for (int i = 0; i < nt; i++)
                for (int w = 0; w < nw; w++) A.gzw[(l * nt + i) * nw + w] /= S.cv / vcorr;
        }
        S.spec_out[0].swap(A.gzw);
        S.spec_out[1].swap(A.jxw);
        S.spec_out[2].swap(A.jyw);
        S.spec_out[3].swap(A.fdq);
    }
    for (int b = 0; b < nb; b++) {
        double vbox = S.sys[b].area * S.thickness;
        for (int k = 0; k < nt; k++) {
            gz[b * nt + k] *= S.Eeff / vbox;
            jx[b * nt + k] *= S.Eeff * 1e30 / vbox;
            jy[b * nt + k] *= S.Eeff * 1e30 / vbox;
        }
        for (int k = 1; k < nt; k++) {
            gz[b * nt + k] += gz[b * nt + k - 1];
            jx[b * nt + k] += jx[b * nt + k - 1];
            jy[b * nt + k] += jy[b * nt + k - 1];
        }
    }
    double vcorr = S.thickness / S.m.lattz;
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < nt; k++) {
            T_out[k * nb + b] = gz[b * nt + k] / (S.cv / vcorr) + S.T0;
            jx_out[k * nb + b] = jx[b * nt + k];
            jy_out[k * nb + b] = jy[b * nt + k];
        }
    if (counters) {
        counters[0] = tot.events;
        counters[1] = tot.crossings;
        counters[2] = tot.nudges;
    }
    return 0;
}

}  // namespace orc
