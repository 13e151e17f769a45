// RTAMC2D transient solver on the GPU: one lane per particle history,
// lanes refill from a global particle counter when their particle ends.
//
// Reference: src/RTA_MC2D.cpp
//   Init_generator (wall part)  272-533   -> init_particle()
//   diffusive_bouncing::bounce  635-678   -> per-(side, mode) alias rows
//   process_segment (transient) 699-769   -> tally_segment()
//   evolparticle (transient)    990-1386  -> body of rta_history_kernel
//   history loop                2228-2317 -> flight draw in the kernel
//   post-processing             2326-2435 -> rta_postprocess()
#include "runtime.hpp"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

namespace bte {

struct RtaGen {
    double p1x, p1y, ndx, ndy;
    int32_t idi, sign, pad0, pad1;
};

struct RtaParams {
    const double4* modes;  // (vx, vy, w, omega)
    const double2* ivel;   // (1/vx, 1/vy), 0 where the component is 0
    const uint2* scat;     // BE_derivative alias (n)
    const DevBox* boxes;
    const DevContact* cts;
    const RtaGen* gens;
    const uint2* gen_sel;  // alias over generators (ngen)
    const uint2* gen_tab;  // ngen x n
    const int4* lam_info;  // per mode: lo, len, base, 0
    const uint2* lam_tab;
    const int32_t* perm;   // omega-sorted position -> mode
    const double* timegrid;
    double* gz;
    double* jx;
    double* jy;
    unsigned long long* counters;  // events, crossings
    unsigned long long* next_pid;
    uint64_t nparticles;
    uint64_t seed;
    unsigned long long max_iters;
    double maxtime;
    int32_t n, nbox, ngen, nt, ballistic;
};

__device__ inline int time_bin(const double* grid, int nt, double target) {
    // get_bin_index (RTA_MC2D.cpp:140-148): lower_bound - 1 clamped to [0, nt-1]
    int lo = 0, hi = nt + 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (grid[mid] < target) lo = mid + 1; else hi = mid;
    }
    int b = lo - 1;
    return b < 0 ? 0 : (b > nt - 1 ? nt - 1 : b);
}

// process_segment (RTA_MC2D.cpp:699-769)
__device__ inline void tally_segment(const RtaParams& P, int box, double t0, double t1, int sign,
                                     double vx, double vy) {
    int b0 = time_bin(P.timegrid, P.nt, t0), b1 = time_bin(P.timegrid, P.nt, t1);
    double delta = t1 - t0;
    for (int i = b0; i <= b1; i++) {
        double lo = fmax(P.timegrid[i] - t0, 0.);
        double hi = fmin(P.timegrid[i + 1] - t0, delta);
        double c = sign * fmax(hi - lo, 0.);
        if (c == 0.) continue;
        size_t k = (size_t)box * P.nt + i;
        atomicAdd(P.gz + k, c);
        atomicAdd(P.jx + k, vx * c);
        atomicAdd(P.jy + k, vy * c);
    }
}

__global__ __launch_bounds__(256) void rta_history_kernel(RtaParams P) {
    Pcg32 rng;
    double x = 0, y = 0, t = 0, dtrem = 0;
    int mode = 0, box = 0, sign = 1;
    bool alive = false, need_flight = true;
    unsigned long long events = 0, crossings = 0, stuck = 0;
    unsigned long long budget = 0;
    int still = 0;        // consecutive events without time advance
    double tlast = -1.0;
    const double maxtime = P.maxtime;
    // particle ids come from a per-wave pool refilled 64 at a time, so the
    // global id counter sees one atomic per 64 histories
    unsigned long long pool_lo = 0, pool_hi = 0;
    for (;;) {
        // every history ends: a particle that exceeds the iteration budget or
        // makes 64 events in a row without advancing in time is dropped and
        // counted, so each lane reaches the exit
        if (alive) {
            still = (t == tlast) ? still + 1 : 0;
            tlast = t;
            if (++budget > P.max_iters || still > 64) {
                stuck++;
                alive = false;
            }
        }
        unsigned long long pid = ~0ULL;
        unsigned long long want = __ballot(!alive);
        while (want) {
            if (pool_lo == pool_hi) {
                unsigned long long base = 0;
                const int leader = __ffsll((long long)want) - 1;
                if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(P.next_pid, 64ULL);
                base = __shfl(base, leader, 64);
                pool_lo = base;
                pool_hi = base + 64;
            }
            const unsigned long long lt = want & ((1ULL << (threadIdx.x & 63)) - 1ULL);
            const unsigned long long rank = (unsigned long long)__popcll(lt);
            const unsigned long long take = min((unsigned long long)__popcll(want), pool_hi - pool_lo);
            const bool mine = (want >> (threadIdx.x & 63)) & 1ULL;
            if (mine && rank < take) pid = pool_lo + rank;
            pool_lo += take;
            want = __ballot(!alive && pid == ~0ULL);
        }
        if (!alive) {
            budget = 0;
            still = 0;
            tlast = -1.0;
            if (pid >= P.nparticles) break;
            rng.seed(P.seed, pid);
            // Init_generator (RTA_MC2D.cpp:436-484)
            bool ng;
            int gi = (int)alias_pick(P.gen_sel, (uint32_t)P.ngen, rng, ng);
            const RtaGen g = P.gens[gi];
            double u = rng.uniform();
            x = g.p1x + u * g.ndx;
            y = g.p1y + u * g.ndy;
            const DevBox bi = P.boxes[g.idi];
            if (!box_inside(bi, x, y)) {
                double dx = 0.5 * (bi.xmin + bi.xmax) - x, dy = 0.5 * (bi.ymin + bi.ymax) - y;
                double dn = sqrt(dx * dx + dy * dy);
                x += 1.0e-6 * dx / dn;
                y += 1.0e-6 * dy / dn;
            }
            mode = (int)alias_pick(P.gen_tab + (size_t)gi * P.n, (uint32_t)P.n, rng, ng);
            sign = g.sign;
            box = g.idi;
            t = 0.;
            alive = true;
            need_flight = true;
        }
        const double4 md = P.modes[mode];
        if (need_flight) {
            // RTA_MC2D.cpp:2262-2288: exponential flight, truncated at maxtime
            if (!(t < maxtime)) { alive = false; continue; }
            double u = 1.0 - rng.uniform();
            dtrem = -log(u) / md.z;
            if (t + dtrem > maxtime) dtrem = maxtime - t + 1.0e-8;
            if (almost_equal(dtrem, 0.)) continue;
            need_flight = false;
        }
        const DevBox B = P.boxes[box];
        if (B.kind == 1) {  // reservoir: absorbed (RTA_MC2D.cpp:1046-1053)
            events++;
            alive = false;
            continue;
        }
        const double vx = md.x, vy = md.y;
        double nx = x + dtrem * vx, ny = y + dtrem * vy;
        if (box_inside(B, nx, ny)) {
            tally_segment(P, box, t, t + dtrem, sign, vx, vy);
            x = nx;
            y = ny;
            t += dtrem;
            if (!P.ballistic) {
                bool ng;
                mode = (int)alias_pick(P.scat, (uint32_t)P.n, rng, ng);
            }
            events++;
            need_flight = true;
            continue;
        }
        int mask;
        double ex, ey;
        const double2 iv = P.ivel[mode];
        double tf = box_exit(B, x, y, vx, vy, iv.x, iv.y, mask, ex, ey);
        double left = dtrem - tf;
        if (almost_equal(left, 0.)) left = 0.;
        tally_segment(P, box, t, t + tf, sign, vx, vy);
        t += tf;
        int side = 0;
        const int fast = side_fast(B, mask, ex, ey, vx, vy, side);
        Ids4 pb{};
        int npb;
        if (fast >= 0) {
            pb.a = fast;
            npb = 1;
        } else {
            npb = fast == SIDE_VOID ? 0 : find_pboxes(P.boxes, P.cts, B, ex, ey, vx, vy, pb);
        }
        if (npb == 0) {
            // void side: Lambert re-emission, remaining flight discarded (RTA_MC2D.cpp:1259)
            int s = fast == SIDE_VOID ? side
                  : (__popc(mask) == 1 ? __ffs(mask) - 1
                                       : pick_void_side(P.boxes, P.cts, B, mask, ex, ey, rng.next()));
            const int4 li = P.lam_info[mode];
            bool ng;
            int k = (int)alias_pick(P.lam_tab + li.z + (size_t)s * li.y, (uint32_t)li.y, rng, ng);
            mode = P.perm[li.x + k];
            x = ex;
            y = ey;
            events++;
            need_flight = true;
            continue;
        }
        x = ex;
        y = ey;
        if (npb == 1) {
            box = pb.a;
        } else {
            // several candidate boxes at a corner (RTA_MC2D.cpp:1316-1380)
            Ids4 cand{};
            int nc = 0;
            double cx = ex + 1.0e-6 * vx, cy = ey + 1.0e-6 * vy;
            for (int k = 0; k < npb && k < 4; k++)
                if (box_inside(P.boxes[pb.get(k)], cx, cy)) cand.set(nc++, pb.get(k));
            if (nc > 0) {
                box = cand.get((int)rng.bounded((uint32_t)nc));
            } else {
                box = pb.get((int)rng.bounded((uint32_t)(npb < 4 ? npb : 4)));
                const DevBox nb = P.boxes[box];
                double dx = 0.5 * (nb.xmin + nb.xmax) - x, dy = 0.5 * (nb.ymin + nb.ymax) - y;
                double dn = sqrt(dx * dx + dy * dy);
                x += 1.0e-6 * dx / dn;
                y += 1.0e-6 * dy / dn;
            }
        }
        crossings++;
        dtrem = left;
        if (dtrem == 0.) need_flight = true;
    }
    atomicAdd(P.counters + 0, events);
    atomicAdd(P.counters + 1, crossings);
    if (stuck) atomicAdd(P.counters + 3, stuck);
}

struct Rta {
    Ctx* ctx = nullptr;
    const Modes* m = nullptr;
    const Geom* g = nullptr;
    bte_rta_cfg cfg{};
    int ntime = 0;
    double maxtime = 0, Eeff = 0, cv = 0, einfo1 = 0;
    std::vector<double> timegrid;
    DevBuf<double4> d_modes;
    DevBuf<double2> d_ivel;
    DevBuf<uint2> d_scat, d_gen_sel, d_gen_tab, d_lam_tab;
    DevBuf<RtaGen> d_gens;
    DevBuf<int4> d_lam_info;
    DevBuf<int32_t> d_perm;
    DevBuf<double> d_timegrid, d_tally;
    DevBuf<unsigned long long> d_counters;
    int ngen = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~Rta() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

static void upload_alias(DevBuf<uint2>& buf, const std::vector<uint32_t>& th,
                         const std::vector<uint32_t>& al, hipStream_t s) {
    std::vector<uint2> v(th.size());
    for (size_t i = 0; i < th.size(); i++) v[i] = make_uint2(th[i], al[i]);
    buf.upload(v, s);
}

// Lambert tables (RTA_MC2D.cpp:556-626 with a constant broadening sigma per
// mode): for mode i the partners are the modes within nsigma*sqrt(2)*sigma in
// omega (a contiguous window of the omega-sorted order); for each wall side s
// the weights are max(-n_s.v_f, 0) * N(omega_i - omega_f; 0, sqrt(2) sigma).
static void build_lambert(const Modes& m, double sigma, std::vector<int4>& info,
                          std::vector<uint2>& tab, std::vector<int32_t>& perm) {
    const double nsig = 2.828427124746190097603377448419396157138;
    int n = m.n;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return m.omega[a] < m.omega[b]; });
    std::vector<int> pos(n);
    for (int k = 0; k < n; k++) pos[perm[k]] = k;
    std::vector<double> sw(n);
    for (int k = 0; k < n; k++) sw[k] = m.omega[perm[k]];
    double sig = std::hypot(sigma, sigma);
    double cut = nsig * sig;
    info.assign(n, make_int4(0, 0, 0, 0));
    tab.clear();
    std::vector<double> pmf;
    std::vector<uint32_t> th, al;
    for (int i = 0; i < n; i++) {
        double oi = m.omega[i];
        int lo = (int)(std::lower_bound(sw.begin(), sw.end(), oi - cut) - sw.begin());
        int hi = (int)(std::upper_bound(sw.begin(), sw.end(), oi + cut) - sw.begin());
        // the float window edges may admit a partner just outside |d| <= cut
        while (lo < hi && std::fabs(oi - sw[lo]) > cut) lo++;
        while (hi > lo && std::fabs(oi - sw[hi - 1]) > cut) hi--;
        int len = hi - lo;
        if (len <= 0) throw Error(BTE_EINVAL, "empty Lambert window");
        info[i] = make_int4(lo, len, (int)tab.size(), 0);
        pmf.resize(len);
        th.resize(len);
        al.resize(len);
        for (int s = 0; s < 4; s++) {
            double nx, ny;
            side_normal(s, nx, ny);
            for (int k = 0; k < len; k++) {
                int f = perm[lo + k];
                double d = std::fabs(oi - m.omega[f]);
                double vn = -(nx * m.vx[f] + ny * m.vy[f]);
                bool ok = !(vn < 0. || almost_equal(vn, 0.)) && !almost_equal(0., m.omega[f]);
                pmf[k] = ok ? vn * std::exp(-0.5 * d * d / (sig * sig)) : 0.;
            }
            if (!build_alias(pmf.data(), len, th.data(), al.data())) {
                // a mode that can reach this side needs a partner leaving it
                if (nx * m.vx[i] + ny * m.vy[i] > 0. && !almost_equal(0., m.omega[i]))
                    throw Error(BTE_EINVAL, "mode " + std::to_string(i) +
                                                " has no diffuse partner within the Lambert window; raise sigma_bounce");
                for (int k = 0; k < len; k++) { th[k] = 0u; al[k] = (uint32_t)(pos[i] - lo); }
            }
            for (int k = 0; k < len; k++) tab.push_back(make_uint2(th[k], al[k]));
        }
    }
}

}  // namespace bte

using namespace bte;

extern "C" {

int bte_rta_create(bte_ctx* ctx, const bte_modes* mm, const bte_geom* gg, const bte_rta_cfg* cfg,
                   bte_rta** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    auto* m = reinterpret_cast<const Modes*>(mm);
    auto* g = reinterpret_cast<const Geom*>(gg);
    if (!c || !m || !g || !cfg || !out) throw Error(BTE_EINVAL, "null argument");
    if (!(cfg->dt > 0) || !(cfg->maxtime > 0) || cfg->nparticles == 0 || !(cfg->thickness > 0))
        throw Error(BTE_EINVAL, "bad RTA configuration");
    HIPCHECK(hipSetDevice(c->device));
    auto* r = new Rta();
    std::unique_ptr<Rta> guard(r);
    r->ctx = c;
    r->m = m;
    r->g = g;
    r->cfg = *cfg;
    const int n = m->n, nbox = (int)g->boxes.size();
    for (auto& b : g->boxes)
        if (b.kind == 0) {
            if (b.Teq != cfg->T0) throw Error(BTE_EINVAL, "Tref in xml is different from provided T0");
            if (b.Tinit != b.Teq) throw Error(BTE_EINVAL, "Initial temperature out of reference is not allowed");
        }
    for (int i = 0; i < n; i++)
        if (m->w[i] == 0.) throw Error(BTE_EINVAL, "invalid scattering rate");
    // time grid (RTA_MC2D.cpp:2121-2124)
    r->ntime = (int)std::floor(cfg->maxtime / cfg->dt + 1.0e-6) + 1;
    if (r->ntime < 2) throw Error(BTE_EINVAL, "maxtime shorter than dt");
    r->maxtime = (r->ntime - 1) * cfg->dt;
    r->timegrid.resize(r->ntime);
    for (int i = 0; i < r->ntime; i++) r->timegrid[i] = (double)i * r->maxtime / (r->ntime - 1);
    hipStream_t s = c->stream;
    // mode records
    std::vector<double4> md(n);
    for (int i = 0; i < n; i++) md[i] = make_double4(m->vx[i], m->vy[i], m->w[i], m->omega[i]);
    r->d_modes.upload(md, s);
    std::vector<double2> iv(n);
    for (int i = 0; i < n; i++)
        iv[i] = make_double2(m->vx[i] != 0. ? 1.0 / m->vx[i] : 0., m->vy[i] != 0. ? 1.0 / m->vy[i] : 0.);
    r->d_ivel.upload(iv, s);
    // BE_derivative sampler (sampling.cpp:47-68)
    std::vector<double> pmf(n);
    for (int i = 0; i < n; i++) pmf[i] = m->omega[i] <= 0. ? 0. : h_be_kernel(m->omega[i], cfg->T0) * m->w[i];
    std::vector<uint32_t> th(n), al(n);
    if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty scattering pmf");
    upload_alias(r->d_scat, th, al, s);
    // isothermal wall generators (RTA_MC2D.cpp:296-347, 366-397)
    double vuc = m->V * cfg->thickness / m->lattz;
    std::vector<RtaGen> gens;
    std::vector<double> gw;
    std::vector<uint2> gtab;
    for (int sid = 0; sid < nbox; sid++) {
        const RectBox& sb = g->boxes[sid];
        if (sb.kind != 1) continue;
        for (auto& hc : g->contacts[sid]) {
            const RectBox& cb = g->boxes[hc.box];
            if (cb.kind != 0) continue;
            if (almost_equal(hc.length(), 0.)) continue;
            if (almost_equal(sb.Teq, cb.Teq)) continue;
            double cum = wall_pmf(*m, sb.Teq, cb.Teq, hc.npx, hc.npy, pmf.data());
            int sg = (sb.Teq - cb.Teq) >= 0 ? 1 : -1;
            double jd = sg * cum / (m->nq * vuc);
            gw.push_back(std::fabs(jd) * hc.length() * cfg->thickness);
            if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty wall pmf");
            for (int i = 0; i < n; i++) gtab.push_back(make_uint2(th[i], al[i]));
            gens.push_back(RtaGen{hc.p1x, hc.p1y, hc.p2x - hc.p1x, hc.p2y - hc.p1y, hc.box, sg, 0, 0});
        }
    }
    if (gens.empty()) throw Error(BTE_EINVAL, "no isothermal generators (reservoirs at a different Teq)");
    r->ngen = (int)gens.size();
    r->einfo1 = std::accumulate(gw.begin(), gw.end(), 0.);
    r->Eeff = r->einfo1 / (double)cfg->nparticles;
    std::vector<uint32_t> gth(gens.size()), gal(gens.size());
    build_alias(gw.data(), (int)gw.size(), gth.data(), gal.data());
    upload_alias(r->d_gen_sel, gth, gal, s);
    r->d_gens.upload(gens, s);
    r->d_gen_tab.upload(gtab, s);
    // Lambert tables
    std::vector<int4> info;
    std::vector<uint2> ltab;
    std::vector<int32_t> perm;
    build_lambert(*m, cfg->sigma_bounce > 0 ? cfg->sigma_bounce : 0.25, info, ltab, perm);
    r->d_lam_info.upload(info, s);
    r->d_lam_tab.upload(ltab, s);
    r->d_perm.upload(perm, s);
    r->d_timegrid.upload(r->timegrid, s);
    r->d_tally.alloc((size_t)3 * nbox * (r->ntime - 1));
    r->d_counters.alloc(4);
    r->cv = h_calc_cv(*m, cfg->T0);
    HIPCHECK(hipEventCreate(&r->ev0));
    HIPCHECK(hipEventCreate(&r->ev1));
    HIPCHECK(hipStreamSynchronize(s));
    *out = reinterpret_cast<bte_rta*>(guard.release());
    BTE_GUARD_END
}

int bte_rta_execute(bte_rta* rr, double* T_out, double* jx_out, double* jy_out, bte_rta_report* rep) {
    BTE_GUARD_BEGIN
    auto* r = reinterpret_cast<Rta*>(rr);
    if (!r || !T_out || !jx_out || !jy_out) throw Error(BTE_EINVAL, "null argument");
    Ctx* c = r->ctx;
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int nbox = (int)r->g->boxes.size(), nt = r->ntime - 1;
    const size_t cells = (size_t)nbox * nt;
    HIPCHECK(hipMemsetAsync(r->d_tally.p, 0, 3 * cells * sizeof(double), s));
    HIPCHECK(hipMemsetAsync(r->d_counters.p, 0, 4 * sizeof(unsigned long long), s));
    RtaParams P{};
    P.modes = r->d_modes.p;
    P.ivel = r->d_ivel.p;
    P.scat = r->d_scat.p;
    P.boxes = r->g->d_boxes.p;
    P.cts = r->g->d_contacts.p;
    P.gens = r->d_gens.p;
    P.gen_sel = r->d_gen_sel.p;
    P.gen_tab = r->d_gen_tab.p;
    P.lam_info = r->d_lam_info.p;
    P.lam_tab = r->d_lam_tab.p;
    P.perm = r->d_perm.p;
    P.timegrid = r->d_timegrid.p;
    P.gz = r->d_tally.p;
    P.jx = r->d_tally.p + cells;
    P.jy = r->d_tally.p + 2 * cells;
    P.counters = r->d_counters.p;
    P.next_pid = r->d_counters.p + 2;
    P.nparticles = r->cfg.nparticles;
    P.seed = r->cfg.seed;
    P.maxtime = r->maxtime;
    P.max_iters = 1ULL << 26;
    P.n = r->m->n;
    P.nbox = nbox;
    P.ngen = r->ngen;
    P.nt = nt;
    P.ballistic = r->cfg.ballistic;
    int blocks = c->cu_count * 8;
    HIPCHECK(hipEventRecord(r->ev0, s));
    hipLaunchKernelGGL(rta_history_kernel, dim3(blocks), dim3(256), 0, s, P);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(r->ev1, s));
    std::vector<double> tally(3 * cells);
    unsigned long long cnt[4];
    HIPCHECK(hipMemcpyAsync(tally.data(), r->d_tally.p, 3 * cells * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(cnt, r->d_counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    // post-processing (RTA_MC2D.cpp:2336-2435)
    double* gz = tally.data();
    double* jx = gz + cells;
    double* jy = jx + cells;
    const double th = r->cfg.thickness;
    for (int b = 0; b < nbox; b++) {
        double vbox = r->g->boxes[b].area() * th;
        for (int k = 0; k < nt; k++) {
            gz[b * nt + k] *= r->Eeff / vbox;
            jx[b * nt + k] *= r->Eeff * 1e30 / vbox;
            jy[b * nt + k] *= r->Eeff * 1e30 / vbox;
        }
        for (int k = 1; k < nt; k++) {
            gz[b * nt + k] += gz[b * nt + k - 1];
            jx[b * nt + k] += jx[b * nt + k - 1];
            jy[b * nt + k] += jy[b * nt + k - 1];
        }
    }
    double vcorr = th / r->m->lattz;
    for (int b = 0; b < nbox; b++)
        for (int k = 0; k < nt; k++) {
            T_out[(size_t)k * nbox + b] = gz[b * nt + k] / (r->cv / vcorr) + r->cfg.T0;
            jx_out[(size_t)k * nbox + b] = jx[b * nt + k];
            jy_out[(size_t)k * nbox + b] = jy[b * nt + k];
        }
    if (cnt[3] != 0)
        throw Error(BTE_EGEOM, std::to_string(cnt[3]) + " particle histories exceeded the iteration budget");
    if (rep) {
        rep->events = cnt[0];
        rep->crossings = cnt[1];
        rep->Eeff = r->Eeff;
        rep->kernel_ms = ms;
        rep->ntime = r->ntime;
    }
    BTE_GUARD_END
}

int bte_rta_destroy(bte_rta* rr) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<Rta*>(rr);
    BTE_GUARD_END
}

int bte_rta_run(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_rta_cfg* cfg,
                double* T_out, double* jx_out, double* jy_out, bte_rta_report* rep) {
    bte_rta* r = nullptr;
    int rc = bte_rta_create(ctx, m, g, cfg, &r);
    if (rc != BTE_OK) return rc;
    rc = bte_rta_execute(r, T_out, jx_out, jy_out, rep);
    bte_rta_destroy(r);
    return rc;
}

}  // extern "C"
