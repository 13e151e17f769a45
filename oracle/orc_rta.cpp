// Oracle (test infrastructure only): restatement of the RTAMC2D transient
// branch (src/RTA_MC2D.cpp) for rectangular geometries and one material.
//   Init_generator          RTA_MC2D.cpp:272-533 (isothermal-wall part)
//   diffusive_bouncing      RTA_MC2D.cpp:547-679
//   process_segment         RTA_MC2D.cpp:699-769, get_bin_index 140-148
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
#include "orc_geometry.hpp"
#include <omp.h>

namespace orc {

struct RtaSetup {
    Modes m;
    std::vector<Box> sys;
    double thickness = 1.0, T0 = 300.0, dt = 1.0, maxtime = 10.0;
    bool ballistic = false;
    double sigma = 0.25;  // Lambert broadening per mode [rad/ps]

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

    void prepare(uint64_t nparticles) {
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
        lambert.assign(m.n, {});
        double sig = std::hypot(sigma, sigma);
        double norm = 1.0 / (sig * std::sqrt(2 * PI));
        for (int i = 0; i < m.n; i++) {
            if (almost_equal(0., m.omega[i])) continue;
            for (int f = 0; f < m.n; f++) {
                if (almost_equal(0., m.omega[f])) continue;
                double d = std::fabs(m.omega[i] - m.omega[f]);
                if (d <= nsigma * sig)
                    lambert[i].push_back({f, norm * std::exp(-0.5 * d * d / (sig * sig))});
            }
        }
    }

    // diffusive_bouncing::bounce (RTA_MC2D.cpp:635-678)
    int bounce(double npx, double npy, int mode, pcg32& rng) const {
        const auto& tm = lambert[mode];
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
};

struct Part {
    double x, y, t;
    int mode, box, sign;
    bool dead;
};

struct Tally {
    int nbox, nt;
    std::vector<double> gz, jx, jy;  // [box][tbin]
    Tally(int nb, int nt_) : nbox(nb), nt(nt_), gz(nb * nt_, 0.), jx(nb * nt_, 0.), jy(nb * nt_, 0.) {}
};

static inline int bin_index(const std::vector<double>& grid, double target) {
    long idx = (long)(std::lower_bound(grid.begin(), grid.end(), target) - grid.begin()) - 1;
    long hi = (long)grid.size() - 2;
    return (int)std::clamp(idx, 0L, hi);
}

// process_segment, transient overload (RTA_MC2D.cpp:699-769)
static void process_segment(const RtaSetup& S, Tally& T, int box, double t0,
                            double t1, int sign, double vx, double vy) {
    int b0 = bin_index(S.timegrid, t0), b1 = bin_index(S.timegrid, t1);
    double delta = t1 - t0;
    for (int i = b0; i <= b1; i++) {
        double lo = std::max(S.timegrid[i] - t0, 0.);
        double hi = std::min(S.timegrid[i + 1] - t0, delta);
        double contr = sign * std::max(hi - lo, 0.);
        T.gz[box * T.nt + i] += contr;
        T.jx[box * T.nt + i] += vx * contr;
        T.jy[box * T.nt + i] += vy * contr;
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
        process_segment(S, T, p.box, p.t, p.t + dt, p.sign, vx, vy);
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
        process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy);
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
        p.mode = S.bounce(ub.npx, ub.npy, p.mode, rng);
        p.x = rx;
        p.y = ry;
        C.events++;
        dt = 0.;
        return;
    }
    if (pboxes.size() == 1) {
        p.x = rx;
        p.y = ry;
        process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy);
        p.t += tf;
        p.box = pboxes[0];
        C.crossings++;
        return;
    }
    process_segment(S, T, p.box, p.t, p.t + tf, p.sign, vx, vy);
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
            double w = S.m.w[p.mode];
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
