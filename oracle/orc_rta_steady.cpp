// Oracle (test infrastructure only): restatement of the RTAMC2D steady
// branch selected by a nonzero <gradient> (src/RTA_MC2D.cpp), for
// rectangular geometries and one material.
//   Nabla_T_distribution        sampling.cpp:118-187
//   Init_generator (gradient)   RTA_MC2D.cpp:347-533
//   evolparticle (steady)       RTA_MC2D.cpp:1406-1799
//   process_segment (steady)    RTA_MC2D.cpp:771-826 (phi on relaxation)
//   generation loop             RTA_MC2D.cpp:2544-2694
//   post-processing, CSV        RTA_MC2D.cpp:2696-2782
// Deviations applied identically on the GPU path: per-particle PCG32 streams
// keyed by (seed, generation, index); a gradient particle's sign alternates
// with its id (even -, odd +) instead of with the global draw order, the
// mode then drawn from the pmf restricted to that sign (the reference
// redraws until the sign matches).

namespace orc {

// Nabla_T_distribution for an in-plane gradient (gx, gy) [K/nm] at T: pmf
// |(-g.v) kernel(omega, T)|, sign of -g.v (>= 0 is +), cumulsum kB sum(pmf)
struct NablaT {
    GridDist plus, minus;
    double cumulsum = 0.;
    bool has_plus = false, has_minus = false;
};
inline NablaT nabla_T(const Modes& m, double gx, double gy, double T) {
    NablaT g;
    std::vector<double> pp(m.n, 0.), pm(m.n, 0.);
    double acc = 0;
    for (int i = 0; i < m.n; i++) {
        if (m.omega[i] == 0.) continue;
        const double vgt = -(gx * m.vx[i] + gy * m.vy[i]);
        const double p = std::fabs(vgt * bose_einstein_kernel(m.omega[i], T));
        acc += p;
        (vgt >= 0. ? pp : pm)[i] = p;
    }
    g.cumulsum = kB * acc;
    double sp = 0, sm = 0;
    for (int i = 0; i < m.n; i++) { sp += pp[i]; sm += pm[i]; }
    if (sp > 0) { g.plus.fill_cumulative(pp); g.has_plus = true; }
    if (sm > 0) { g.minus.fill_cumulative(pm); g.has_minus = true; }
    return g;
}

struct SPart {
    double x, y;
    int mode, box, sign;
    uint64_t id;
};

struct SteadySetup {
    RtaSetup base;            // material, geometry, walls, Lambert, sampler
    double gx = 0., gy = 0.;  // <gradient x y> [K/nm]
    double eps_energy = -42., eps_flux = -42.;  // RTA_MC2D.cpp:210-211
    int max_generations = 100000;
    std::vector<int> gbox;          // boxes with a gradient generator
    std::vector<NablaT> ggen;
    std::vector<double> gcum;       // cumulative box weights
    double Ewall = 0., Egrad = 0.;  // ExGen
};

// steady evolparticle (RTA_MC2D.cpp:1406-1799): one flight; a relaxation
// inside a box adds the sign to phi(box); a diffuse wall bounce passes the
// particle (new mode, at the wall) to the next generation
static void steady_evol(const SteadySetup& SS, SPart& p, double vx, double vy, double& dt, pcg32& rng,
                        std::vector<double>& gz, std::vector<double>& jx, std::vector<double>& jy,
                        std::vector<double>& phi, std::vector<SPart>& next, bool& dead) {
    const RtaSetup& S = SS.base;
    const auto& sys = S.sys;
    const Box& box = sys[p.box];
    auto seg = [&](int b, double t) {
        gz[b] += p.sign * t;
        jx[b] += vx * p.sign * t;
        jy[b] += vy * p.sign * t;
    };
    if (box.periodic) {
        p.box = translate(sys, p.box, p.x, p.y, vx, vy, rng);
        return;
    }
    if (box.reservoir) {
        dead = true;
        dt = 0.;
        return;
    }
    double nx = p.x + dt * vx, ny = p.y + dt * vy;
    if (box.inside(nx, ny)) {
        p.x = nx;
        p.y = ny;
        seg(p.box, dt);
        phi[p.box] += p.sign;
        dead = true;
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
        return;
    }
    const double tf = hit.time;
    double left = dt - tf;
    if (almost_equal(left, 0.)) left = 0.;
    dt = left;
    double rx = hit.rx, ry = hit.ry;
    const double vn = std::hypot(vx, vy);
    const double ex = rx + 1.0e-6 * vx / vn, ey = ry + 1.0e-6 * vy / vn;
    std::vector<int> pboxes;
    for (auto& [c, bc] : box.contacts) {
        if (c == p.box) continue;
        const double d = bc.npx * vx + bc.npy * vy;
        if ((sys[c].inside(rx, ry) || sys[c].inside(ex, ey)) && !almost_equal(d, 0.) && d > 0.) pboxes.push_back(c);
    }
    if (pboxes.empty()) {
        seg(p.box, tf);
        std::vector<int> cids;
        std::vector<int> bids = hit.bids;
        if (in_corner3(rx, ry, sys, cids)) {
            auto cs = correct_corner_problem(rx, ry, vx, vy, sys, cids, rng);
            const int bid = cs.first;
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
        next.push_back(p);
        dead = true;
        dt = 0.;
        return;
    }
    seg(p.box, tf);
    p.x = rx;
    p.y = ry;
    if (pboxes.size() == 1) {
        p.box = pboxes[0];
        return;
    }
    std::vector<int> pbox2;
    const double cx = rx + 1.0e-6 * vx, cy = ry + 1.0e-6 * vy;
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
}

static inline uint64_t gen_key(uint64_t seed, int gen) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(gen + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Outputs per box: T, jx, jy (the three rows of steady_<T0>K_run_<i>.csv).
int rta_steady_run(SteadySetup& SS, uint64_t nparticles, uint64_t seed, int nthreads, double* T_out,
                   double* jx_out, double* jy_out, int* generations, double* Eeff_out) {
    RtaSetup& S = SS.base;
    S.prepare(nparticles);  // walls, Lambert table, sampler, cv
    auto& sys = S.sys;
    const int nb = (int)sys.size();
    const double vuc = S.m.V * S.thickness / S.m.lattz;
    // gradient generators (RTA_MC2D.cpp:354-367, 402-427)
    SS.gbox.clear();
    SS.ggen.clear();
    std::vector<double> gw;
    if (!almost_equal(std::hypot(SS.gx, SS.gy), 0.)) {
        for (auto& s : sys) {
            if (s.reservoir || s.periodic) continue;
            NablaT g = nabla_T(S.m, SS.gx, SS.gy, s.Teq);
            SS.gbox.push_back(s.id);
            gw.push_back((s.area * S.thickness / vuc) * g.cumulsum / S.m.nq);
            SS.ggen.push_back(std::move(g));
        }
    }
    SS.Ewall = S.einfo1;
    SS.Egrad = 0;
    for (double w : gw) SS.Egrad += w;
    SS.gcum.clear();
    {
        double c = 0;
        for (double w : gw) { c += w / SS.Egrad; SS.gcum.push_back(c); }
    }
    const double etot = SS.Ewall + SS.Egrad;
    if (!(etot > 0)) throw value_error("no wall or gradient generators");
    S.Eeff = etot / (double)nparticles;
    const double pwall = SS.Ewall / etot;
    // initial particles (Init_generator)
    std::vector<SPart> parts(nparticles);
    for (uint64_t i = 0; i < nparticles; i++) {
        pcg32 rng(seed, i);
        const double r = rng.uniform();
        SPart q{};
        q.id = i;
        if (SS.Ewall > 0 && r <= pwall) {
            Part p = init_particle(S, rng);
            q.x = p.x; q.y = p.y; q.mode = p.mode; q.box = p.box; q.sign = p.sign;
        } else {
            const double u = rng.uniform();
            size_t k = std::lower_bound(SS.gcum.begin(), SS.gcum.end(), u) - SS.gcum.begin();
            if (k >= SS.gbox.size()) k = SS.gbox.size() - 1;
            const Box& b = sys[SS.gbox[k]];
            std::tie(q.x, q.y) = b.random_point(rng);
            q.box = b.id;
            const NablaT& g = SS.ggen[k];
            int sg = (i & 1) ? 1 : -1;
            if (sg > 0 && !g.has_plus) sg = -1;
            if (sg < 0 && !g.has_minus) sg = 1;
            q.sign = sg;
            q.mode = (int)(sg > 0 ? g.plus : g.minus).sample(rng);
        }
        parts[i] = q;
    }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nth = omp_get_max_threads();
    std::vector<double> gz(nb, 0.), jx(nb, 0.), jy(nb, 0.);
    bool converged = false;
    int gen = 0;
    std::string err;
    while (!parts.empty() && !converged && gen < SS.max_generations) {
        std::vector<std::vector<double>> tgz(nth, std::vector<double>(nb, 0.)), tjx = tgz, tjy = tgz, tphi = tgz;
        std::vector<std::vector<SPart>> tnext(nth);
        const uint64_t key = gen_key(seed, gen);
#pragma omp parallel for schedule(dynamic, 64)
        for (long long i = 0; i < (long long)parts.size(); i++) {
            const int tid = omp_get_thread_num();
            try {
                pcg32 rng(key, (uint64_t)i);
                SPart p = parts[i];
                double vx, vy;
                get_v(S.m, p.mode, vx, vy);
                const double w = S.m.w[p.mode];
                if (w == 0.) throw value_error("invalid scattering rate");
                double dt = -std::log(1.0 - rng.uniform()) / w;
                bool dead = false;
                long guard = 0;
                while (!almost_equal(dt, 0.) && !dead) {
                    steady_evol(SS, p, vx, vy, dt, rng, tgz[tid], tjx[tid], tjy[tid], tphi[tid], tnext[tid], dead);
                    if (++guard > 10000000) throw geometry_error("runaway flight");
                }
            } catch (const std::exception& e) {
#pragma omp critical
                err = e.what();
            }
        }
        if (!err.empty()) throw geometry_error(err);
        std::vector<double> ogz = gz, ojx = jx, ojy = jy, phi(nb, 0.);
        std::vector<SPart> next;
        for (int t = 0; t < nth; t++) {
            for (int b = 0; b < nb; b++) {
                gz[b] += tgz[t][b];
                jx[b] += tjx[t][b];
                jy[b] += tjy[t][b];
                phi[b] += tphi[t][b];
            }
            next.insert(next.end(), tnext[t].begin(), tnext[t].end());
        }
        // bounced particles in id order (independent of thread scheduling)
        std::sort(next.begin(), next.end(), [](const SPart& a, const SPart& b) { return a.id < b.id; });
        // re-emission of the relaxed energy: |phi| particles per box
        // (RTA_MC2D.cpp:2637-2654)
        uint64_t nid = 0;
        for (auto& q : next) q.id = nid++;
        for (int b = 0; b < nb; b++) {
            if (almost_equal(phi[b], 0.)) continue;
            const int sg = phi[b] >= 0 ? 1 : -1;
            const long cnt = (long)std::fabs(phi[b]);
            pcg32 rng(gen_key(seed ^ 0x5eedULL, gen), (uint64_t)b);
            for (long k = 0; k < cnt; k++) {
                SPart q{};
                std::tie(q.x, q.y) = sys[b].random_point(rng);
                q.mode = (int)S.sampler.sample(rng);
                q.box = b;
                q.sign = sg;
                q.id = nid++;
                next.push_back(q);
            }
        }
        auto rel = [](const std::vector<double>& o, const std::vector<double>& n) {
            double d = 0, s = 0;
            for (size_t k = 0; k < n.size(); k++) { d += (o[k] - n[k]) * (o[k] - n[k]); s += n[k] * n[k]; }
            return std::sqrt(d) / std::sqrt(s);
        };
        const double rgz = rel(ogz, gz), rjx = rel(ojx, jx), rjy = rel(ojy, jy);
        if (almost_equal(SS.gx, 0.)) converged = rgz < SS.eps_energy && rjy < SS.eps_flux;
        else if (almost_equal(SS.gy, 0.)) converged = rgz < SS.eps_energy && rjx < SS.eps_flux;
        else converged = rgz < SS.eps_energy && rjx < SS.eps_flux && rjy < SS.eps_flux;
        parts.swap(next);
        gen++;
    }
    // post-processing (RTA_MC2D.cpp:2696-2728)
    const double vcorr = S.thickness / S.m.lattz;
    for (int b = 0; b < nb; b++) {
        const double vbox = sys[b].area * S.thickness;
        T_out[b] = gz[b] / vbox * S.Eeff / (S.cv / vcorr) + S.T0;
        jx_out[b] = jx[b] / vbox * S.Eeff * 1e30;
        jy_out[b] = jy[b] / vbox * S.Eeff * 1e30;
    }
    if (generations) *generations = gen;
    if (Eeff_out) *Eeff_out = S.Eeff;
    return 0;
}

}  // namespace orc
