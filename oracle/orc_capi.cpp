// Oracle C interface (test infrastructure only; loaded with ctypes by the
// tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg).
//
// Parity pinning of each oracle function (what checks it in tests/):
//   pcg32                       reference KAT external/pcg-cpp-0.98/test-high/
//                               expected/check-pcg32.out (all 5 rounds)
//   bose_einstein(_kernel),
//   calc_cv, wall flux,
//   calc_Eeff                   closed forms (numpy) + limits
//   GridDist sampling           exact pmf on small N (chi^2) and brute force
//   CubicSpline1D               scipy CubicSpline(bc_type='natural')
//   expm                        scipy.linalg.expm golden fixtures, 2x2 closed
//                               form, diagonal matrices
//   arnoldi / build_P           m >= N exactness, textbook-vs-reference Krylov
//                               relation A V = V H + f e_m^T, expm fixtures
//   synthetic_B                 invariants 1^T B = 0, B C = 0, B offdiag >= 0
//   get_inter_side / contacts   analytic slab intersections on rectangles
//   propagator_H                exact enumeration of R on a grid
//   lirp / propagator selection closed forms of the interpolation line and
//                               the key pair Scattering picks
//   calculate_gradientT         exact gradients of linear Teq fields,
//                               minimum-norm rank-1 case, pseudo-inverse
//                               (numpy.linalg.lstsq) on random layouts
//   source_term_gradient        mean generated energy per step = the
//                               Nabla_T energy dt vol cumulsum/nq (Eeff units)
//   rta_run                     ballistic closed form (no scattering), flux
//                               antisymmetry, energy bookkeeping
//   bert_run                    conservation of the signed count (P >= 0,
//                               adiabatic walls), E[H_{n+1}] = P H_n with v = 0
//   lambert_table / bounce      Lambert cosine law per side (chi-square),
//                               energy window, the odds exp(-d^2/(4 sigma^2))
//                               of a partner at frequency distance d, omega = 0
//                               modes never reached (test_oracle_bounce.py)
//   corner resolution, nudges   ballistic flights through shared corners and
//                               edges end at the exact endpoint (1e-5 nm);
//                               symmetric corners split 1/2 : 1/2
#include "orc_geometry.hpp"
#include <omp.h>

namespace orc {
void expm(int n, const double* A, double* E);
long double error_bound(int N, const double* A, int n, double t);
int krylov_size(int N, const double* A, double t, double eps);
void arnoldi(int N, const double* A, const double* b, int n, double t, int mode, double* V, double* H);
void build_P(int N, const double* A, double t, int m, int mode, double* P);
void build_P_cols(int N, const double* A, double t, int m, int mode, int c0, int c1, double* P);
void project_columns(int N, double* B);
void synthetic_B(int N, const double* C, const double* w, double density, uint64_t seed,
                 int sinkhorn_iters, double* B);
}  // namespace orc

// full definitions are needed for the setup structs
#include "orc_rta.cpp"
#include "orc_bert.cpp"
#include "orc_rta_steady.cpp"

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ---------------- PCG32 ----------------
void orc_pcg32_init(uint64_t initstate, uint64_t initseq, uint64_t* st) {
    pcg32 r(initstate, initseq);
    st[0] = r.state;
    st[1] = r.inc;
}
void orc_pcg32_next(uint64_t* st, int n, uint32_t* out) {
    pcg32 r;
    r.state = st[0];
    r.inc = st[1];
    for (int i = 0; i < n; i++) out[i] = r.next();
    st[0] = r.state;
}
void orc_pcg32_bounded(uint64_t* st, uint32_t bound, int n, uint32_t* out) {
    pcg32 r;
    r.state = st[0];
    r.inc = st[1];
    for (int i = 0; i < n; i++) out[i] = r.bounded(bound);
    st[0] = r.state;
}
void orc_pcg32_uniform(uint64_t* st, int n, double* out) {
    pcg32 r;
    r.state = st[0];
    r.inc = st[1];
    for (int i = 0; i < n; i++) out[i] = r.uniform();
    st[0] = r.state;
}

// ---------------- synthetic mode table ----------------
// f ~ U(fmin,fmax) THz, omega = 2 pi f; |v| ~ U(vmin,vmax) nm/ps with angle
// ~ U(0, 2pi); tau = 10^U(log10 tmin, log10 tmax) ps, w = 1/tau. With
// mirror != 0 the second half repeats the first with vy -> -vy.
// Draw order per mode: f, |v|, angle, tau from PCG32(seed, stream 0x6d6f6465).
void orc_synthetic_modes(int n, uint64_t seed, double fmin, double fmax, double vmin,
                         double vmax, double tmin, double tmax, int mirror, double* omega,
                         double* vx, double* vy, double* w) {
    pcg32 r(seed, 0x6d6f6465u);
    int nbase = mirror ? n / 2 : n;
    double l0 = std::log10(tmin), l1 = std::log10(tmax);
    for (int i = 0; i < nbase; i++) {
        double f = fmin + (fmax - fmin) * r.uniform();
        double sp = vmin + (vmax - vmin) * r.uniform();
        double th = 2.0 * PI * r.uniform();
        double tau = std::pow(10.0, l0 + (l1 - l0) * r.uniform());
        omega[i] = 2.0 * PI * f;
        vx[i] = sp * std::cos(th);
        vy[i] = sp * std::sin(th);
        w[i] = 1.0 / tau;
    }
    if (mirror)
        for (int i = 0; i < nbase; i++) {
            omega[nbase + i] = omega[i];
            vx[nbase + i] = vx[i];
            vy[nbase + i] = -vy[i];
            w[nbase + i] = w[i];
        }
}

// ---------------- closed-form physics ----------------
void orc_bose_einstein(int n, const double* omega, double T, double* out) {
    for (int i = 0; i < n; i++) out[i] = bose_einstein(omega[i], T);
}
void orc_bose_einstein_kernel(int n, const double* omega, double T, double* out) {
    for (int i = 0; i < n; i++) out[i] = bose_einstein_kernel(omega[i], T);
}
static Modes mk_modes(int n, const double* omega, const double* vx, const double* vy,
                      const double* w, double V, double lattz, int nq) {
    Modes m;
    m.n = n;
    m.nq = nq;
    m.omega.assign(omega, omega + n);
    m.vx.assign(vx, vx + n);
    m.vy.assign(vy, vy + n);
    m.w.assign(w, w + n);
    m.V = V;
    m.lattz = lattz;
    return m;
}
double orc_calc_cv(int n, const double* omega, double V, int nq, double T) {
    Modes m;
    m.n = n;
    m.nq = nq;
    m.V = V;
    m.omega.assign(omega, omega + n);
    return calc_cv(m, T);
}
// wall distribution: cumulative (n) and flux for vuc; returns cumulsum
double orc_wall_dist(int n, const double* omega, const double* vx, const double* vy, int nq,
                     double Twall, double Teq, double nx, double ny, double vuc,
                     double* cumulative, double* flux) {
    Modes m;
    m.n = n;
    m.nq = nq;
    m.omega.assign(omega, omega + n);
    m.vx.assign(vx, vx + n);
    m.vy.assign(vy, vy + n);
    m.w.assign(n, 1.0);
    GridDist d = isothermal_wall(m, Twall, Teq, nx, ny);
    if (cumulative) std::memcpy(cumulative, d.cumulative.data(), n * sizeof(double));
    if (flux) *flux = wall_flux(d, vuc);
    return d.cumulsum;
}
void orc_spline(int n, const double* X, const double* Y, int nq, const double* q, double* out) {
    CubicSpline1D s(std::vector<double>(X, X + n), std::vector<double>(Y, Y + n));
    for (int i = 0; i < nq; i++) out[i] = s.interpolate(q[i]);
}
// GridDist sampling: draw ns samples from the pmf p (n) with PCG32(seed, stream)
void orc_grid_sample(int n, const double* p, uint64_t seed, uint64_t stream, int ns, int32_t* out) {
    GridDist d;
    d.fill_cumulative(std::vector<double>(p, p + n));
    pcg32 r(seed, stream);
    for (int i = 0; i < ns; i++) out[i] = (int32_t)d.sample(r);
}
// propagator_H::get for a column b at the given R values
void orc_propH_get(int n, const double* b, int nr, const double* R, int32_t* sign, int32_t* idx) {
    PropH h(b, n);
    for (int i = 0; i < nr; i++) {
        auto s = h.get(R[i]);
        sign[i] = s.first;
        idx[i] = s.second;
    }
}

// The propagator used at temperature Tx from those at Ts (ntemp column-major
// ncol x ncol matrices in Ps, ascending Ts): the tabulated one, lirp of the
// signed cumulatives (rule 0) or the interpolated matrix (rule 1); the
// signed cumulative of every column into sc (column-major). Returns 0, or 1
// when Tx cannot be interpolated.
int orc_prop_interp(int ncol, int ntemp, const double* Ts, const double* Ps, double Tx, int rule, double* sc) {
    try {
        std::vector<double> keys(Ts, Ts + ntemp);
        auto col = [&](int t, int c) { return PropH(Ps + ((size_t)t * ncol + c) * ncol, ncol); };
        int exact = -1;
        for (int t = 0; t < ntemp; t++)
            if (Ts[t] == Tx) exact = t;
        std::pair<double, double> pr{0., 0.};
        int t1 = 0, t2 = 0;
        if (exact < 0) {
            pr = prop_pair(keys, Tx, rule);
            t1 = (int)(std::find(keys.begin(), keys.end(), pr.first) - keys.begin());
            t2 = (int)(std::find(keys.begin(), keys.end(), pr.second) - keys.begin());
        }
        std::vector<double> Pm;
        if (exact < 0 && rule == 1) {
            std::vector<double> P1(Ps + (size_t)t1 * ncol * ncol, Ps + (size_t)(t1 + 1) * ncol * ncol);
            std::vector<double> P2(Ps + (size_t)t2 * ncol * ncol, Ps + (size_t)(t2 + 1) * ncol * ncol);
            Pm = interp_matrix(P1, P2, (Tx - pr.first) / (pr.second - pr.first));
        }
        for (int c = 0; c < ncol; c++) {
            PropH h = exact >= 0 ? col(exact, c)
                      : (rule == 1 ? PropH(Pm.data() + (size_t)c * ncol, ncol)
                                   : lirp(col(t1, c), col(t2, c), pr.first, pr.second, Tx));
            std::copy(h.signed_cumm.begin(), h.signed_cumm.end(), sc + (size_t)c * ncol);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
// propagator_H::get on a given signed cumulative (e.g. from lirp)
void orc_propH_get_sc(int n, const double* sc, int nr, const double* R, int32_t* sign, int32_t* idx) {
    PropH h(std::vector<double>(sc, sc + n), 1.0);
    for (int i = 0; i < nr; i++) {
        auto s = h.get(R[i]);
        sign[i] = s.first;
        idx[i] = s.second;
    }
}

// ---------------- geometry ----------------
// periodic maps for the next run (target box and translation per box),
// set by orc_set_periodic and consumed by mk_sys
static thread_local std::vector<int32_t> g_ptarget;
static thread_local std::vector<double> g_ptrans;
// convex polygon boxes for the next run (vertex counts and x, y pairs), set
// by orc_set_polygons and consumed by mk_sys in place of the rectangles
static thread_local std::vector<int32_t> g_nvert;
static thread_local std::vector<double> g_xy;

static std::vector<Box> mk_sys(int nbox, const double* rect, const int32_t* kind,
                               const double* Teq, const double* Tinit) {
    std::vector<Box> sys(nbox);
    if (!g_nvert.empty() && (int)g_nvert.size() != nbox) throw geometry_error("polygon count differs from boxes");
    size_t voff = 0;
    for (int i = 0; i < nbox; i++) {
        sys[i].id = i;
        if (g_nvert.empty()) {
            sys[i].make(rect[4 * i + 0], rect[4 * i + 1], rect[4 * i + 2], rect[4 * i + 3]);
        } else {
            std::vector<double> vx(g_nvert[i]), vy(g_nvert[i]);
            for (int k = 0; k < g_nvert[i]; k++) {
                vx[k] = g_xy[2 * (voff + k)];
                vy[k] = g_xy[2 * (voff + k) + 1];
            }
            voff += g_nvert[i];
            sys[i].make_poly(vx, vy);
        }
        sys[i].reservoir = kind[i] == 1;
        sys[i].periodic = kind[i] == 2;
        sys[i].Teq = Teq[i];
        sys[i].Treal = Tinit ? Tinit[i] : Teq[i];
        if (sys[i].periodic) {
            if ((int)g_ptarget.size() != nbox) throw geometry_error("periodic box without a translation");
            sys[i].target = g_ptarget[i];
            sys[i].tx = g_ptrans[2 * i];
            sys[i].ty = g_ptrans[2 * i + 1];
        }
    }
    g_ptarget.clear();
    g_ptrans.clear();
    g_nvert.clear();
    g_xy.clear();
    calculate_contacts(sys);
    return sys;
}
void orc_set_polygons(int nbox, const int32_t* nvert, const double* xy) {
    g_nvert.assign(nvert, nvert + nbox);
    size_t nv = 0;
    for (int i = 0; i < nbox; i++) nv += (size_t)nvert[i];
    g_xy.assign(xy, xy + 2 * nv);
}
void orc_set_periodic(int nbox, const int32_t* target, const double* trans) {
    g_ptarget.assign(target, target + nbox);
    g_ptrans.assign(trans, trans + 2 * (size_t)nbox);
}
// calculate_gradientT: (gx, gy) per box into grad[2*nbox]
int orc_gradient_T(int nbox, const double* rect, const int32_t* kind, const double* Teq, double* grad) {
    try {
        auto sys = mk_sys(nbox, rect, kind, Teq, nullptr);
        auto g = gradient_T(sys);
        for (int i = 0; i < nbox; i++) {
            grad[2 * i] = g[i][0];
            grad[2 * i + 1] = g[i][1];
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
// get_inter_side on box ibox: returns 0 ok, 1 geometry_error; out[0]=time,
// out[1..2]=point, bids (up to 4) with nbids
int orc_inter_side(int nbox, const double* rect, int ibox, double x0, double y0, double vx,
                   double vy, double dt, double* out, int32_t* bids, int32_t* nbids) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        auto h = sys[ibox].get_inter_side(x0, y0, vx, vy, dt);
        out[0] = h.time;
        out[1] = h.rx;
        out[2] = h.ry;
        *nbids = (int32_t)h.bids.size();
        for (size_t k = 0; k < h.bids.size() && k < 4; k++) bids[k] = h.bids[k];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
// contacts: for each box, list (neighbour, npx, npy, p1x, p1y, p2x, p2y)
int orc_box_info(int nbox, const double* rect, double* out) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        for (int i = 0; i < nbox; i++) {
            double* o = out + 8 * (size_t)i;
            o[0] = sys[i].area; o[1] = sys[i].cx; o[2] = sys[i].cy;
            o[3] = sys[i].xmin; o[4] = sys[i].xmax; o[5] = sys[i].ymin; o[6] = sys[i].ymax;
            o[7] = (double)sys[i].borders.size();
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_borders(int nbox, const double* rect, int ibox, int maxb, double* out) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        int k = 0;
        for (auto& b : sys[ibox].borders) {
            if (k >= maxb) break;
            double* o = out + 6 * (size_t)k++;
            o[0] = b.p1x; o[1] = b.p1y; o[2] = b.p2x; o[3] = b.p2y; o[4] = b.npx; o[5] = b.npy;
        }
        return k;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int orc_inside(int nbox, const double* rect, int ibox, int npts, const double* xy, int32_t* out) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        for (int k = 0; k < npts; k++) out[k] = sys[ibox].inside(xy[2 * k], xy[2 * k + 1]) ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_random_points(int nbox, const double* rect, int ibox, int n, uint64_t seed, double* xy) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        pcg32 rng(seed, 7u);
        for (int k = 0; k < n; k++) std::tie(xy[2 * k], xy[2 * k + 1]) = sys[ibox].random_point(rng);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_contacts(int nbox, const double* rect, int maxc, int32_t* ncont, int32_t* nbr, double* geo) {
    std::vector<int32_t> kind(nbox, 0);
    std::vector<double> T(nbox, 300.);
    try {
        auto sys = mk_sys(nbox, rect, kind.data(), T.data(), nullptr);
        for (int i = 0; i < nbox; i++) {
            int k = 0;
            for (auto& [c, b] : sys[i].contacts) {
                if (k >= maxc) break;
                nbr[i * maxc + k] = c;
                double* g = geo + (size_t)(i * maxc + k) * 6;
                g[0] = b.npx; g[1] = b.npy; g[2] = b.p1x; g[3] = b.p1y; g[4] = b.p2x; g[5] = b.p2y;
                k++;
            }
            ncont[i] = k;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// ---------------- RTAMC2D ----------------
// kind: 0 bin, 1 reservoir, 2 periodic. Outputs [ntime-1][nbox].
// spectral decomposition of the next orc_rta_run (boxes, ticks, per-mode
// broadening or NULL), and its results after the run
static thread_local std::vector<int> g_spec_boxes;
static thread_local std::vector<double> g_spec_sigma, g_spec_grid, g_spec_out[4];
static thread_local int g_spec_nomega = 0, g_spec_nt = 0, g_spec_nq = 0, g_spec_nw = 0;
void orc_rta_spectral(int nloc, const int32_t* boxes, int nomega, int n, const double* sigma) {
    g_spec_boxes.assign(boxes, boxes + nloc);
    g_spec_nomega = nomega;
    if (sigma) g_spec_sigma.assign(sigma, sigma + n);
    else g_spec_sigma.clear();
}
// location k of the last run: omega[nw], dT/jx/jy [nt][nw], fd [nt][nq]
int orc_rta_spectral_get(int k, double* omega, double* dT, double* jx, double* jy, double* fd) {
    if (g_spec_out[0].empty() || k < 0 || (size_t)k * g_spec_nt * g_spec_nw >= g_spec_out[0].size()) return 1;
    const size_t nw = (size_t)g_spec_nt * g_spec_nw, nq = (size_t)g_spec_nt * g_spec_nq;
    std::copy(g_spec_grid.begin(), g_spec_grid.end(), omega);
    std::copy(g_spec_out[0].begin() + k * nw, g_spec_out[0].begin() + (k + 1) * nw, dT);
    std::copy(g_spec_out[1].begin() + k * nw, g_spec_out[1].begin() + (k + 1) * nw, jx);
    std::copy(g_spec_out[2].begin() + k * nw, g_spec_out[2].begin() + (k + 1) * nw, jy);
    std::copy(g_spec_out[3].begin() + k * nq, g_spec_out[3].begin() + (k + 1) * nq, fd);
    return 0;
}
double orc_gamma_pdf(double k, double theta, double x) { return gamma_pdf(k, theta, x); }
// extensions of the next orc_rta_run (no reference counterpart): scattering
// rate factor, specular probability per side (left, bottom, top, right) and
// the sides whose dT/dp is tallied; the derivative after the run
static thread_local double g_ext_wscale = 1.0, g_ext_spec[4] = {0., 0., 0., 0.};
static thread_local int g_ext_gmask = 0;
static thread_local std::vector<double> g_grad_out;
void orc_rta_ext(double wscale, const double* spec4, int gmask) {
    g_ext_wscale = wscale;
    for (int k = 0; k < 4; k++) g_ext_spec[k] = spec4 ? spec4[k] : 0.;
    g_ext_gmask = gmask;
}
// per-mode base_sigma of the Lambert bounce of the next orc_rta_run (NULL:
// the constant sigma)
static thread_local std::vector<double> g_ext_msig;
void orc_rta_mode_sigma(int n, const double* sigma) {
    if (sigma) g_ext_msig.assign(sigma, sigma + n);
    else g_ext_msig.clear();
}
// dT/dp of the last run: [4][nt][nbox]
int orc_rta_grad_get(double* out) {
    if (g_grad_out.empty()) return 1;
    std::copy(g_grad_out.begin(), g_grad_out.end(), out);
    return 0;
}
int orc_rta_run(int n, const double* omega, const double* vx, const double* vy, const double* w,
                double V, double lattz, int nq, int nbox, const double* rect, const int32_t* kind,
                const double* Teq, double thickness, double T0, double dt, double maxtime,
                uint64_t nparticles, uint64_t seed, int ballistic, double sigma, int nthreads,
                double* T_out, double* jx_out, double* jy_out, uint64_t* counters, double* raw_gz,
                double* Eeff_out) {
    try {
        RtaSetup S;
        S.m = mk_modes(n, omega, vx, vy, w, V, lattz, nq);
        S.sys = mk_sys(nbox, rect, kind, Teq, nullptr);
        S.thickness = thickness;
        S.T0 = T0;
        S.dt = dt;
        S.maxtime = maxtime;
        S.ballistic = ballistic != 0;
        S.sigma = sigma;
        S.wscale = g_ext_wscale;
        for (int k = 0; k < 4; k++) S.spec[k] = g_ext_spec[k];
        S.gmask = g_ext_gmask;
        S.mode_sigma.swap(g_ext_msig);
        g_ext_msig.clear();
        if (!S.mode_sigma.empty() && (int)S.mode_sigma.size() != n) throw input_error("per-mode sigma of the wrong size");
        g_ext_wscale = 1.0;
        for (int k = 0; k < 4; k++) g_ext_spec[k] = 0.;
        g_ext_gmask = 0;
        if (!(S.wscale > 0.)) throw input_error("wscale must be positive");
        S.spec_boxes.swap(g_spec_boxes);
        S.spec_sigma.swap(g_spec_sigma);
        S.nomega = g_spec_nomega;
        g_spec_boxes.clear();
        g_spec_sigma.clear();
        g_spec_nomega = 0;
        rta_run(S, nparticles, seed, nthreads, T_out, jx_out, jy_out, counters, raw_gz);
        if (Eeff_out) *Eeff_out = S.Eeff;
        g_spec_grid = S.omegagrid;
        for (int k = 0; k < 4; k++) g_spec_out[k].swap(S.spec_out[k]);
        g_spec_nt = S.ntime - 1;
        g_spec_nq = S.m.nq;
        g_spec_nw = S.nomega;
        g_grad_out.swap(S.grad_out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// diffusive_bouncing alone: ns bounces of `mode` at a side with outward
// normal (npx, npy), stream k of `seed` for bounce k; out: the new modes
// (msig: per-mode base_sigma or NULL for sigma on every mode)
int orc_lambert_sample(int n, const double* omega, const double* vx, const double* vy, double sigma, double npx,
                       double npy, int mode, uint64_t seed, int ns, int32_t* out, const double* msig) {
    try {
        std::vector<double> w(n, 1.0);
        Modes m = mk_modes(n, omega, vx, vy, w.data(), 1.0, 1.0, 1);
        const LambertTable L = lambert_table(m, sigma, msig ? std::vector<double>(msig, msig + n) : std::vector<double>());
        for (int k = 0; k < ns; k++) {
            pcg32 rng(seed, (uint64_t)k);
            out[k] = lambert_bounce(m, L, npx, npy, mode, rng);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// One ballistic evolparticle event (RTA_MC2D.cpp:990-1386) of a particle at
// (x, y) in box ibox with `mode` and flight time dt, repeated ns times with
// stream k of `seed`: out_box, out_xy (2 per repeat), out_mode, out_dt (the
// flight left) and out_kind (0 flight ended inside, 1 wall bounce, 2 box
// crossing, 3 nudge, 4 absorbed, 5 periodic translation).
int orc_rta_event(int n, const double* omega, const double* vx, const double* vy, double sigma, int nbox,
                  const double* rect, const int32_t* kind, int ibox, double x, double y, int mode, double dt,
                  uint64_t seed, int ns, int32_t* out_box, double* out_xy, int32_t* out_mode, double* out_dt,
                  int32_t* out_kind) {
    try {
        std::vector<double> w(n, 1.0), Teq(nbox, 300.0);
        RtaSetup S;
        S.m = mk_modes(n, omega, vx, vy, w.data(), 1.0, 1.0, 1);
        S.sys = mk_sys(nbox, rect, kind, Teq.data(), nullptr);
        S.ballistic = true;
        S.sigma = sigma;
        S.lambert = lambert_table(S.m, sigma, {});
        S.timegrid = {0.0, 1.0e30};
        S.ntime = 2;
        if (ibox < 0 || ibox >= nbox || mode < 0 || mode >= n) throw input_error("bad particle");
        for (int k = 0; k < ns; k++) {
            pcg32 rng(seed, (uint64_t)k);
            Tally T(nbox, 1);
            Counters C;
            Part p{x, y, 0., mode, ibox, 1, false};
            double v_x, v_y;
            get_v(S.m, p.mode, v_x, v_y);
            double d = dt;
            evolparticle(S, p, v_x, v_y, d, rng, T, C);
            out_box[k] = p.box;
            out_xy[2 * k] = p.x;
            out_xy[2 * k + 1] = p.y;
            out_mode[k] = p.mode;
            out_dt[k] = d;
            const bool inside = S.sys[ibox].inside(x + dt * v_x, y + dt * v_y);
            out_kind[k] = p.dead ? 4 : C.nudges ? 3 : C.crossings ? 2 : C.events ? (inside ? 0 : 1) : 5;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// ---------------- beRTAMC2D ----------------
// steady branch (nonzero gradient): per box T, jx, jy
int orc_rta_steady_run(int n, const double* omega, const double* vx, const double* vy, const double* w,
                       double V, double lattz, int nq, int nbox, const double* rect, const int32_t* kind,
                       const double* Teq, double thickness, double T0, double gx, double gy, double eps_energy,
                       double eps_flux, int max_generations, uint64_t nparticles, uint64_t seed, double sigma,
                       int nthreads, double* T_out, double* jx_out, double* jy_out, int32_t* generations,
                       double* Eeff_out) {
    try {
        SteadySetup SS;
        RtaSetup& S = SS.base;
        S.m = mk_modes(n, omega, vx, vy, w, V, lattz, nq);
        S.sys = mk_sys(nbox, rect, kind, Teq, nullptr);
        S.thickness = thickness;
        S.T0 = T0;
        S.dt = 1.0;
        S.maxtime = 1.0;
        S.sigma = sigma;
        SS.gx = gx;
        SS.gy = gy;
        SS.eps_energy = eps_energy;
        SS.eps_flux = eps_flux;
        if (max_generations > 0) SS.max_generations = max_generations;
        int gens = 0;
        rta_steady_run(SS, nparticles, seed, nthreads, T_out, jx_out, jy_out, &gens, Eeff_out);
        *generations = gens;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
// propagators at several temperatures, the pair rule and a fixed gradient
// for the next orc_bert_run (Ps: ntemp column-major matrices, any T order)
static thread_local std::vector<double> g_bT;
static thread_local std::vector<std::vector<double>> g_bP;
static thread_local int g_brule = 0, g_bfixed = 0, g_bcancel = 4;
static thread_local double g_bgx = 0., g_bgy = 0.;
// given initial distributions of the next orc_bert_run: ninit boxes, each
// with len entries (the modes after the Gamma offset)
static thread_local std::map<int, std::vector<double>> g_binit;
void orc_bert_init_dist(int ninit, const int32_t* boxes, int len, const double* fd) {
    g_binit.clear();
    for (int k = 0; k < ninit; k++) g_binit[boxes[k]].assign(fd + (size_t)k * len, fd + (size_t)(k + 1) * len);
}
void orc_bert_options(int ncol, int ntemp, const double* Ts, const double* Ps, int rule, int fixed_gradient,
                      double gx, double gy, int pcancel_each) {
    std::vector<int> ord(ntemp);
    for (int t = 0; t < ntemp; t++) ord[t] = t;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return Ts[a] < Ts[b]; });
    g_bT.clear();
    g_bP.clear();
    for (int t : ord) {
        g_bT.push_back(Ts[t]);
        g_bP.emplace_back(Ps + (size_t)t * ncol * ncol, Ps + (size_t)(t + 1) * ncol * ncol);
    }
    g_brule = rule;
    g_bfixed = fixed_gradient;
    g_bgx = gx;
    g_bgy = gy;
    g_bcancel = pcancel_each > 0 ? pcancel_each : 4;
}
int orc_bert_run(int n, const double* omega, const double* vx, const double* vy, const double* w,
                 double V, double lattz, int nq, int nbox, const double* rect, const int32_t* kind,
                 const double* Teq, const double* Tinit, double thickness, double dt,
                 double maxtime, int ngamma, const double* P, int rta, double Ed,
                 uint64_t Efactor, double Tmax, double Tmin, uint64_t seed, int nthreads,
                 int max_steps, double* props, int64_t* npart, int32_t* H_last, double* Eeff_out,
                 int32_t* nsteps_out, int64_t* pending) {
    try {
        BertSetup S;
        S.m = mk_modes(n, omega, vx, vy, w, V, lattz, nq);
        S.sys = mk_sys(nbox, rect, kind, Teq, Tinit);
        S.thickness = thickness;
        S.dt = dt;
        S.maxtime = maxtime;
        S.ngamma = ngamma;
        S.rta = rta != 0;
        if (!S.rta) S.P.assign(P, P + (size_t)(n - ngamma) * (n - ngamma));
        S.Ed = Ed;
        S.Efactor = Efactor;
        S.Tmax = Tmax;
        S.Tmin = Tmin;
        S.PT.swap(g_bT);
        S.Ps.swap(g_bP);
        S.prop_rule = g_brule;
        S.fixed_gradient = g_bfixed != 0;
        S.gx = g_bgx;
        S.gy = g_bgy;
        S.pcancel_each = g_bcancel;
        S.init_dist.swap(g_binit);
        g_binit.clear();
        g_bT.clear();
        g_bP.clear();
        g_brule = g_bfixed = 0;
        g_bcancel = 4;
        g_bgx = g_bgy = 0.;
        *nsteps_out = bert_run(S, seed, nthreads, max_steps, props, npart, H_last, Eeff_out, pending);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// ---------------- Krylov ----------------
void orc_expm(int n, const double* A, double* E) { expm(n, A, E); }
double orc_error_bound(int N, const double* A, int n, double t) { return (double)error_bound(N, A, n, t); }
int orc_krylov_size(int N, const double* A, double t, double eps) { return krylov_size(N, A, t, eps); }
void orc_arnoldi(int N, const double* A, const double* b, int n, double t, int mode, double* V,
                 double* H) {
    arnoldi(N, A, b, n, t, mode, V, H);
}
void orc_build_P(int N, const double* A, double t, int m, int mode, int nthreads, double* P) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    build_P(N, A, t, m, mode, P);
}
void orc_build_P_cols(int N, const double* A, double t, int m, int mode, int c0, int c1, int nthreads, double* P) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    build_P_cols(N, A, t, m, mode, c0, c1, P);
}
void orc_project_columns(int N, double* B) { project_columns(N, B); }
void orc_synthetic_B(int N, const double* C, const double* w, double density, uint64_t seed,
                     int iters, double* B) {
    synthetic_B(N, C, w, density, seed, iters, B);
}

}  // extern "C"
