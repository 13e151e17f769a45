// Host side of the C ABI: errors, context, material and geometry handles,
// synthetic inputs, alias tables and the small host-side physics used to
// build device tables.
#include "runtime.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <limits>

namespace bte {

static thread_local std::string g_last_error;
void set_error(const std::string& s) { g_last_error = s; }

double h_bose_einstein(double omega, double T) {
    constexpr double pf = 1e12 * hbar / kB;
    return 1. / (std::exp(pf * omega / T) - 1.);
}
// include/utilities.hpp:401-415
double h_be_kernel(double omega, double T) {
    constexpr double pf = 0.5 * 1e12 * hbar / kB;
    double x = pf * omega / T;
    if (x < 1e-320) return 1.;
    if (x > 300.) return 0.;
    double s = x / std::sinh(x);
    return s * s;
}
// include/bulk_properties.hpp:32-47
double h_calc_cv(const Modes& m, double T) {
    double acc = 0;
    for (int i = 0; i < m.n; i++) acc += h_be_kernel(m.omega[i], T);
    return acc * kB / m.nq / m.V;
}

bool build_alias(const double* pmf, int n, uint32_t* thresh, uint32_t* alias) {
    double tot = 0;
    for (int i = 0; i < n; i++) {
        if (!(pmf[i] >= 0.)) throw Error(BTE_EINVAL, "probabilities must be >= 0.");
        tot += pmf[i];
    }
    if (!(tot > 0.)) {
        for (int i = 0; i < n; i++) { thresh[i] = 0xffffffffu; alias[i] = (uint32_t)i; }
        return false;
    }
    std::vector<double> q(n);
    std::vector<int> small, large;
    small.reserve(n);
    large.reserve(n);
    for (int i = 0; i < n; i++) {
        q[i] = pmf[i] * n / tot;
        (q[i] < 1.0 ? small : large).push_back(i);
    }
    for (int i = 0; i < n; i++) alias[i] = (uint32_t)i;
    while (!small.empty() && !large.empty()) {
        int s = small.back();
        small.pop_back();
        int l = large.back();
        double qs = q[s];
        double t = std::ldexp(qs, 32);
        thresh[s] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)std::llround(t);
        alias[s] = (uint32_t)l;
        q[l] = (q[l] + qs) - 1.0;
        if (q[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (int l : large) { thresh[l] = 0xffffffffu; alias[l] = (uint32_t)l; }
    for (int s : small) { thresh[s] = 0xffffffffu; alias[s] = (uint32_t)s; }
    return true;
}

// R maps to the first index (in std::lower_bound's search) with |sc| >= R,
// a non-decreasing step function of R that only changes at the values |sc|.
// A monotone |sc| gives the increments directly; otherwise every interval
// between consecutive distinct values is mapped through std::lower_bound
// itself, so a non-monotone cumulative (interpolated propagators) is sampled
// exactly as the search would.
void pick_pmf(const double* sc, int n, double* pmf) {
    for (int k = 0; k < n; k++) pmf[k] = 0.;
    if (n <= 0) return;
    bool mono = true;
    for (int k = 1; k < n && mono; k++) mono = std::fabs(sc[k]) >= std::fabs(sc[k - 1]);
    double lo = 0.;
    if (mono) {
        for (int k = 0; k < n; k++) {
            const double hi = std::min(std::fabs(sc[k]), 1.0);
            if (hi > lo) {
                pmf[k] = hi - lo;
                lo = hi;
            }
        }
    } else {
        std::vector<double> v(n);
        for (int k = 0; k < n; k++) v[k] = std::min(std::fabs(sc[k]), 1.0);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        auto abs_less = [](double a, double b) { return std::fabs(a) < std::fabs(b); };
        for (double hi : v) {
            if (!(hi > lo)) continue;
            // every R in (lo, hi] gives the same search result as R = hi
            size_t r = (size_t)(std::lower_bound(sc, sc + n, hi, abs_less) - sc);
            if (r >= (size_t)n) r = (size_t)n - 1;
            pmf[r] += hi - lo;
            lo = hi;
        }
    }
    if (lo < 1.0) pmf[n - 1] += 1.0 - lo;  // R above every |sc|
}

// Minimum-norm least squares h_i . g = Teq_i - Teq_0 over the regular
// contact boxes i of a regular box (h_i: centre differences), through the
// singular value decomposition of the 2-column system; singular values below
// 2 eps sigma_max count as zero.
std::vector<double2> gradient_T(const Geom& g) {
    const int nbox = (int)g.boxes.size();
    std::vector<double2> out(nbox, make_double2(0., 0.));
    for (int i = 0; i < nbox; i++) {
        const RectBox& b0 = g.boxes[i];
        if (b0.kind != 0) continue;
        std::vector<double> hx, hy, df;
        for (const HostContact& hc : g.contacts[i]) {
            const RectBox& bc = g.boxes[hc.box];
            if (bc.kind != 0) continue;
            df.push_back(bc.Teq - b0.Teq);
            hx.push_back(bc.cx() - b0.cx());
            hy.push_back(bc.cy() - b0.cy());
        }
        const size_t m = df.size();
        if (m == 0) continue;
        double a = 0, bb = 0, c = 0;
        for (size_t k = 0; k < m; k++) {
            a += hx[k] * hx[k];
            bb += hy[k] * hy[k];
            c += hx[k] * hy[k];
        }
        // right singular vectors (cs, sn) and (-sn, cs)
        const double th = 0.5 * std::atan2(2. * c, a - bb);
        const double cs = std::cos(th), sn = std::sin(th);
        double s1 = 0, s2 = 0, p1 = 0, p2 = 0;
        for (size_t k = 0; k < m; k++) {
            const double u1 = cs * hx[k] + sn * hy[k];
            const double u2 = -sn * hx[k] + cs * hy[k];
            s1 += u1 * u1;
            s2 += u2 * u2;
            p1 += u1 * df[k];
            p2 += u2 * df[k];
        }
        const double smax = std::sqrt(std::max(s1, s2));
        const double tol = 2. * 2.220446049250313e-16 * (double)std::min<size_t>(m, 2) * smax;
        double gx = 0, gy = 0;
        if (std::sqrt(s1) > tol) { gx += cs * p1 / s1; gy += sn * p1 / s1; }
        if (std::sqrt(s2) > tol) { gx += -sn * p2 / s2; gy += cs * p2 / s2; }
        out[i] = make_double2(gx, gy);
    }
    return out;
}

void spectral_weights(const Modes& m, const double* sigma, double sigma_default, int nomega, const double* grid,
                      std::vector<double>& W) {
    const int n = m.n, nq = m.nq, nb = n / nq, ncols = 3 * nomega + nq;
    // broadening per mode with the low-outlier floor exp(p25 - 1.5 (p75 -
    // p25)) of log sigma (RTA_MC2D.cpp:2034-2037, dist_reader.cpp:327-331,
    // isotopic_scattering.cpp:31-50)
    std::vector<double> sg(n), ls(n);
    for (int i = 0; i < n; i++) {
        sg[i] = sigma ? sigma[i] : sigma_default;
        ls[i] = sg[i] == 0. ? std::numeric_limits<double>::lowest() : std::log(sg[i]);
    }
    std::sort(ls.begin(), ls.end());
    const double p0 = ls[25 * (size_t)n / 100], p1 = ls[75 * (size_t)n / 100];
    const double lb = std::exp(p0 - 1.5 * (p1 - p0));
    W.assign((size_t)n * ncols, 0.);
    for (int i = 0; i < n; i++) {
        const double om = m.omega[i];
        if (almost_equal(om, 0.)) continue;  // Gamma modes carry no particles
        double* row = W.data() + (size_t)i * ncols;
        row[3 * nomega + i / nb] = 1.0 / om;
        const double sgi = std::max(sg[i], lb);
        if (almost_equal(sgi, 0.)) continue;  // skipped as in dist_reader.cpp:357-358
        const double k = om * om / sgi, th = sgi / om, lgk = std::lgamma(k);
        for (int j = 0; j < nomega; j++) {
            const double z = grid[j] / th;
            const double w = z > 0. ? std::exp((k - 1.) * std::log(z) - z - lgk) / th : 0.;
            row[j] = w;
            row[nomega + j] = m.vx[i] * w;
            row[2 * nomega + j] = m.vy[i] * w;
        }
    }
}

// Isothermal_wall_distribution pmf (sampling.cpp:189-269); normal points into
// the box the particles are emitted into.
double wall_pmf(const Modes& m, double Twall, double Teq, double nx, double ny, double* pmf) {
    double acc = 0;
    for (int i = 0; i < m.n; i++) {
        double vn = nx * m.vx[i] + ny * m.vy[i];
        if (vn <= 0. || almost_equal(0., m.omega[i]))
            pmf[i] = 0.;
        else
            pmf[i] = vn * m.omega[i] *
                     std::fabs(h_bose_einstein(m.omega[i], Twall) - h_bose_einstein(m.omega[i], Teq));
        acc += pmf[i];
    }
    return 1.0e12 * hbar * acc;
}

void Spline::build(const std::vector<double>& X, const std::vector<double>& Y) {
    int n = (int)Y.size() - 1;
    std::vector<double> A(Y), B(n), D(n), h(n), alpha(n, 0.), C(n + 1), l(n + 1), mu(n + 1), z(n + 1);
    for (int i = 0; i < n; i++) h[i] = X[i + 1] - X[i];
    for (int i = 1; i < n; i++)
        alpha[i] = 3. / h[i] * (A[i + 1] - A[i]) - 3. / h[i - 1] * (A[i] - A[i - 1]);
    l[0] = 1; mu[0] = 0; z[0] = 0;
    for (int i = 1; i < n; i++) {
        l[i] = 2 * (X[i + 1] - X[i - 1]) - h[i - 1] * mu[i - 1];
        mu[i] = h[i] / l[i];
        z[i] = (alpha[i] - h[i - 1] * z[i - 1]) / l[i];
    }
    l[n] = 1; z[n] = 0; C[n] = 0;
    for (int j = n - 1; j >= 0; j--) {
        C[j] = z[j] - mu[j] * C[j + 1];
        B[j] = (A[j + 1] - A[j]) / h[j] - h[j] * (C[j + 1] + 2 * C[j]) / 3;
        D[j] = (C[j + 1] - C[j]) / (3 * h[j]);
    }
    a.assign(A.begin(), A.begin() + n);
    b = B;
    c.assign(C.begin(), C.begin() + n);
    d = D;
    x0.assign(X.begin(), X.begin() + n);
    xe.assign(X.begin() + 1, X.end());
}
double Spline::eval(double x) const {
    size_t idx = std::lower_bound(xe.begin(), xe.end(), x) - xe.begin();
    if (idx >= a.size()) idx--;
    double dx = x - x0[idx];
    return a[idx] + dx * (b[idx] + dx * (c[idx] + d[idx] * dx));
}

}  // namespace bte

using namespace bte;

extern "C" {

int bte_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return BTE_EINVAL;
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = 0;
    return BTE_OK;
}

int bte_version(void) { return 100; }

int bte_ctx_create(int device, bte_ctx** out) {
    BTE_GUARD_BEGIN
    if (!out) throw Error(BTE_EINVAL, "null output");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error(BTE_ENODEV, "no HIP device visible");
    if (device < 0 || device >= ndev) throw Error(BTE_EINVAL, "device index out of range");
    HIPCHECK(hipSetDevice(device));
    auto* c = new Ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = reinterpret_cast<bte_ctx*>(c);
    BTE_GUARD_END
}

int bte_ctx_destroy(bte_ctx* ctx) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return BTE_OK;
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->blas && c->blas_free) c->blas_free(c->blas);
    HIPCHECK(hipStreamDestroy(c->stream));
    delete c;
    BTE_GUARD_END
}

int bte_sync(bte_ctx* ctx) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    HIPCHECK(hipStreamSynchronize(c->stream));
    BTE_GUARD_END
}

int bte_synthetic_modes(int n, uint64_t seed, double fmin, double fmax, double vmin, double vmax,
                        double tmin, double tmax, int mirror, double* omega, double* vx,
                        double* vy, double* w) {
    BTE_GUARD_BEGIN
    if (n <= 0 || !omega || !vx || !vy || !w) throw Error(BTE_EINVAL, "bad arguments");
    if (mirror && (n % 2)) throw Error(BTE_EINVAL, "mirror tables need an even mode count");
    Pcg32 r;
    r.seed(seed, 0x6d6f6465u);
    const double PI = 3.14159265358979323846;
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
    BTE_GUARD_END
}

int bte_synthetic_B(int N, const double* C, const double* w, double density, uint64_t seed,
                    int iters, double* B) {
    BTE_GUARD_BEGIN
    if (N <= 1 || !C || !w || !B || !(density > 0. && density <= 1.))
        throw Error(BTE_EINVAL, "bad arguments");
    Pcg32 rng;
    rng.seed(seed, 0x6f6du);
    std::vector<int> ri, cj;
    std::vector<double> val;
    uint32_t thr = (uint32_t)(density * 4294967296.0);
    for (int j = 1; j < N; j++)
        for (int i = 0; i < j; i++) {
            uint32_t r = rng.next();
            if (r < thr) {
                ri.push_back(i);
                cj.push_back(j);
                val.push_back((rng.next() >> 8) * (1.0 / 16777216.0) + 1.0 / 33554432.0);
            }
        }
    std::vector<double> d(N, 1.0), rs(N);
    for (int it = 0; it < iters; it++) {
        std::fill(rs.begin(), rs.end(), 0.);
        for (size_t e = 0; e < val.size(); e++) {
            rs[ri[e]] += val[e] * d[cj[e]];
            rs[cj[e]] += val[e] * d[ri[e]];
        }
        // stop at the last finite scaling when the pattern admits no exact one
        std::vector<double> nd(N);
        bool ok = true;
        for (int i = 0; i < N; i++) {
            nd[i] = std::sqrt(d[i] * C[i] * w[i] / (rs[i] > 0 ? rs[i] : 1e-300));
            if (!(nd[i] > 0. && nd[i] < 1e300)) ok = false;
        }
        if (!ok) break;
        d.swap(nd);
    }
    std::fill(B, B + (size_t)N * N, 0.);
    std::vector<double> srow(N, 0.);
    for (size_t e = 0; e < val.size(); e++) {
        double s = d[ri[e]] * val[e] * d[cj[e]];
        B[ri[e] + (size_t)cj[e] * N] = s / C[cj[e]];
        B[cj[e] + (size_t)ri[e] * N] = s / C[ri[e]];
        srow[ri[e]] += s;
        srow[cj[e]] += s;
    }
    for (int j = 0; j < N; j++) B[j + (size_t)j * N] = -srow[j] / C[j];
    BTE_GUARD_END
}

int bte_alias_build(int n, const double* pmf, uint32_t* thresh, uint32_t* alias) {
    BTE_GUARD_BEGIN
    if (n <= 0 || !pmf || !thresh || !alias) throw Error(BTE_EINVAL, "bad arguments");
    if (!build_alias(pmf, n, thresh, alias)) throw Error(BTE_EINVAL, "pmf sums to zero");
    BTE_GUARD_END
}

int bte_modes_create(bte_ctx* ctx, int n, int nq, const double* omega, const double* vx,
                     const double* vy, const double* w, double V, double lattz, bte_modes** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !out || n <= 0 || nq <= 0 || n % nq || !omega || !vx || !vy || !w || !(V > 0) ||
        !(lattz > 0))
        throw Error(BTE_EINVAL, "bad arguments to bte_modes_create");
    if (n >= (1 << 30)) throw Error(BTE_EINVAL, "too many modes");
    auto* m = new Modes();
    m->ctx = c;
    m->n = n;
    m->nq = nq;
    m->omega.assign(omega, omega + n);
    m->vx.assign(vx, vx + n);
    m->vy.assign(vy, vy + n);
    m->w.assign(w, w + n);
    // get_v cleaning (RTA_MC2D.cpp:843-856): components almost_equal to 0 are 0
    for (int i = 0; i < n; i++) {
        if (almost_equal(m->vx[i], 0.)) m->vx[i] = 0.;
        if (almost_equal(m->vy[i], 0.)) m->vy[i] = 0.;
    }
    m->V = V;
    m->lattz = lattz;
    HIPCHECK(hipSetDevice(c->device));
    m->d_vx.upload(m->vx, c->stream);
    m->d_vy.upload(m->vy, c->stream);
    m->d_w.upload(m->w, c->stream);
    m->d_omega.upload(m->omega, c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
    *out = reinterpret_cast<bte_modes*>(m);
    BTE_GUARD_END
}

int bte_modes_set_sigma(bte_modes* mm, const double* sigma) {
    BTE_GUARD_BEGIN
    auto* m = reinterpret_cast<Modes*>(mm);
    if (!m) throw Error(BTE_EINVAL, "null mode table");
    if (!sigma) {
        m->sigma.clear();
        return BTE_OK;
    }
    for (int i = 0; i < m->n; i++)
        if (!(sigma[i] >= 0.) || !std::isfinite(sigma[i]))
            throw Error(BTE_EINVAL, "per-mode sigma must be finite and nonnegative (mode " + std::to_string(i) + ")");
    m->sigma.assign(sigma, sigma + m->n);
    BTE_GUARD_END
}

int bte_modes_destroy(bte_modes* mm) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<Modes*>(mm);
    BTE_GUARD_END
}

namespace {

double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

// Andrew's monotone chain: hull vertices counter-clockwise, collinear and
// repeated points dropped
std::vector<double2> convex_hull(std::vector<double2> p) {
    std::sort(p.begin(), p.end(), [](const double2& a, const double2& b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
    p.erase(std::unique(p.begin(), p.end(), [](const double2& a, const double2& b) { return a.x == b.x && a.y == b.y; }),
            p.end());
    if (p.size() < 3) return {};
    std::vector<double2> h(2 * p.size());
    size_t k = 0;
    for (size_t i = 0; i < p.size(); i++) {
        while (k >= 2 && cross2(h[k - 1].x - h[k - 2].x, h[k - 1].y - h[k - 2].y, p[i].x - h[k - 2].x,
                                p[i].y - h[k - 2].y) <= 0.)
            k--;
        h[k++] = p[i];
    }
    for (size_t i = p.size() - 1, t = k + 1; i-- > 0;) {
        while (k >= t && cross2(h[k - 1].x - h[k - 2].x, h[k - 1].y - h[k - 2].y, p[i].x - h[k - 2].x,
                                p[i].y - h[k - 2].y) <= 0.)
            k--;
        h[k++] = p[i];
    }
    h.resize(k - 1);
    return h;
}

// p within tol of the closed segment a-b
bool near_segment(const double2& p, const double2& a, const double2& b, double tol) {
    const double dx = b.x - a.x, dy = b.y - a.y, L = std::hypot(dx, dy);
    if (L == 0.) return std::hypot(p.x - a.x, p.y - a.y) <= tol;
    if (std::fabs(cross2(dx, dy, p.x - a.x, p.y - a.y)) / L > tol) return false;
    const double u = ((p.x - a.x) * dx + (p.y - a.y) * dy) / L;
    return u >= -tol && u <= L + tol;
}

std::vector<double2> box_vertices(const RectBox& b) {
    if (!b.hull.empty()) return b.hull;
    return {make_double2(b.xmin, b.ymin), make_double2(b.xmax, b.ymin), make_double2(b.xmax, b.ymax),
            make_double2(b.xmin, b.ymax)};
}

// calculate_contacts (geometry_2d.cpp:295-381) for convex polygons: the
// boundary points two boxes share (a vertex of one on the other's
// boundary), a segment between the two farthest of them or a single point;
// the border normal points away from the owner's centre. Boxes whose
// interiors overlap (no separating edge normal) are an error.
bool poly_contact(const RectBox& a, const RectBox& e, double tol, HostContact& hc) {
    const std::vector<double2> A = box_vertices(a), E = box_vertices(e);
    // separating-axis test over the edge normals of both boxes: apart (no
    // contact), touching along some axis, or overlapping interiors
    bool touch = false;
    for (int pass = 0; pass < 2 && !touch; pass++) {
        const std::vector<double2>& P = pass ? E : A;
        for (size_t k = 0; k < P.size() && !touch; k++) {
            const double2 p1 = P[k], p2 = P[(k + 1) % P.size()];
            double nx = p2.y - p1.y, ny = -(p2.x - p1.x);
            const double nn = std::hypot(nx, ny);
            nx /= nn;
            ny /= nn;
            double amin = 1e300, amax = -1e300, emin = 1e300, emax = -1e300;
            for (auto& v : A) { const double d = v.x * nx + v.y * ny; amin = std::min(amin, d); amax = std::max(amax, d); }
            for (auto& v : E) { const double d = v.x * nx + v.y * ny; emin = std::min(emin, d); emax = std::max(emax, d); }
            if (amax < emin - tol || emax < amin - tol) return false;
            touch = amax <= emin + tol || emax <= amin + tol;
        }
    }
    if (!touch) throw Error(BTE_EGEOM, "overlapping boxes");
    std::vector<double2> pts;
    auto add = [&](const double2& q) {
        for (auto& r : pts)
            if (std::hypot(q.x - r.x, q.y - r.y) <= tol) return;
        pts.push_back(q);
    };
    for (auto& v : A)
        for (size_t k = 0; k < E.size(); k++)
            if (near_segment(v, E[k], E[(k + 1) % E.size()], tol)) { add(v); break; }
    for (auto& v : E)
        for (size_t k = 0; k < A.size(); k++)
            if (near_segment(v, A[k], A[(k + 1) % A.size()], tol)) { add(v); break; }
    if (pts.empty()) return false;
    size_t i0 = 0, i1 = 0;
    double best = -1.;
    for (size_t i = 0; i < pts.size(); i++)
        for (size_t j = i + 1; j < pts.size(); j++) {
            const double d = std::hypot(pts[i].x - pts[j].x, pts[i].y - pts[j].y);
            if (d > best) { best = d; i0 = i; i1 = j; }
        }
    hc.p1x = pts[i0].x; hc.p1y = pts[i0].y; hc.p2x = pts[i1].x; hc.p2y = pts[i1].y;
    if (pts.size() == 1 || best < 1.0e-6) {
        const double px = pts.size() == 1 ? pts[0].x : 0.5 * (pts[i0].x + pts[i1].x);
        const double py = pts.size() == 1 ? pts[0].y : 0.5 * (pts[i0].y + pts[i1].y);
        hc.p1x = hc.p2x = px;
        hc.p1y = hc.p2y = py;
        const double nx = px - a.cx(), ny = py - a.cy(), nn = std::hypot(nx, ny);
        hc.npx = nx / nn;
        hc.npy = ny / nn;
        return true;
    }
    const double ndx = hc.p2x - hc.p1x, ndy = hc.p2y - hc.p1y;
    const double mx = 0.5 * (hc.p1x + hc.p2x) - a.cx(), my = 0.5 * (hc.p1y + hc.p2y) - a.cy();
    double nx = ndy, ny = -ndx;
    if (nx * mx + ny * my < 0.) { nx = -nx; ny = -ny; }
    const double nn = std::hypot(nx, ny);
    hc.npx = nx / nn;
    hc.npy = ny / nn;
    return true;
}
}  // namespace

// contacts of rectangles (geometry_2d.cpp:295-381): overlap of the closed
// boxes
static void rect_contacts(Geom* g) {
    const int nbox = (int)g->boxes.size();
    for (int i = 0; i < nbox; i++) {
        const RectBox& a = g->boxes[i];
        for (int j = 0; j < nbox; j++) {
            if (j == i) continue;
            const RectBox& e = g->boxes[j];
            double x0 = std::max(a.xmin, e.xmin), x1 = std::min(a.xmax, e.xmax);
            double y0 = std::max(a.ymin, e.ymin), y1 = std::min(a.ymax, e.ymax);
            if (x0 > x1 || y0 > y1) continue;
            HostContact hc{j, 0, 0, x0, y0, x1, y1};
            if (x0 == x1 && y0 == y1) {
                double nx = x0 - a.cx(), ny = y0 - a.cy(), nn = std::hypot(nx, ny);
                hc.npx = nx / nn;
                hc.npy = ny / nn;
            } else if (x0 != x1 && y0 != y1) {
                throw Error(BTE_EGEOM, "overlapping boxes");
            } else {
                double ndx = x1 - x0, ndy = y1 - y0;
                double mx = 0.5 * (x0 + x1) - a.cx(), my = 0.5 * (y0 + y1) - a.cy();
                double nx = ndy, ny = -ndx;
                if (nx * mx + ny * my < 0.) { nx = -nx; ny = -ny; }
                double nn = std::hypot(nx, ny);
                hc.npx = nx / nn;
                hc.npy = ny / nn;
            }
            g->contacts[i].push_back(hc);
        }
    }
}

// Device tables of a geometry whose boxes and contacts are set: boxes with
// their contact ranges and side neighbours, contact segments, and the sides
// (rectangles: left, bottom, top, right; polygons: hull edges) with the ids
// of their outward normals.
static void finish_geom(Geom* g) {
    const int nbox = (int)g->boxes.size();
    for (int i = 0; i < nbox; i++)
        if (g->contacts[i].empty() && nbox > 1) throw Error(BTE_EGEOM, "There are unconnected boxes");
    double scale = 0.;
    for (auto& b : g->boxes) scale = std::max({scale, std::fabs(b.xmin), std::fabs(b.xmax), std::fabs(b.ymin), std::fabs(b.ymax)});
    const double tol = 1.0e-9 * std::max(scale, 1.0);
    auto normal_id = [&](double nx, double ny) {
        for (size_t k = 0; k < g->normals.size(); k++)
            if (std::hypot(g->normals[k].x - nx, g->normals[k].y - ny) < 1.0e-9) return (int32_t)k;
        g->normals.push_back(make_double2(nx, ny));
        return (int32_t)g->normals.size() - 1;
    };
    if (!g->poly)
        for (int k = 0; k < 4; k++) {
            double nx, ny;
            side_normal(k, nx, ny);
            normal_id(nx, ny);
        }
    g->dboxes.resize(nbox);
    for (int i = 0; i < nbox; i++) {
        DevBox& d = g->dboxes[i];
        const RectBox& b = g->boxes[i];
        d.xmin = b.xmin; d.xmax = b.xmax; d.ymin = b.ymin; d.ymax = b.ymax;
        d.kind = b.kind;
        d.c0 = (int32_t)g->dcontacts.size();
        for (auto& hc : g->contacts[i]) {
            g->dcontacts.push_back(DevContact{hc.box, 0, hc.npx, hc.npy});
            // the contact border as a segment extended by 1e-6 at both ends
            const double L = hc.length();
            const double ux = L > 0 ? (hc.p2x - hc.p1x) / L : 0., uy = L > 0 ? (hc.p2y - hc.p1y) / L : 0.;
            g->cseg.push_back(make_double4(hc.p1x - 1.0e-6 * ux, hc.p1y - 1.0e-6 * uy, hc.p2x + 1.0e-6 * ux,
                                           hc.p2y + 1.0e-6 * uy));
        }
        d.c1 = (int32_t)g->dcontacts.size();
        d.tgt = -1;
        // the sides, and for each the box sharing all of it: VOID when no
        // contact segment of positive length lies on it, MIXED when several
        // or one not covering it
        std::vector<DevSide> sd;
        if (b.hull.empty()) {
            for (int k = 0; k < 4; k++) {
                DevSide q{};
                q.p1x = k == 3 ? b.xmax : b.xmin;
                q.p1y = k == 2 || k == 3 ? b.ymax : b.ymin;
                q.p2x = k == 0 ? b.xmin : b.xmax;
                q.p2y = k == 0 || k == 2 ? b.ymax : b.ymin;
                side_normal(k, q.nx, q.ny);
                sd.push_back(q);
            }
        } else {
            const size_t m = b.hull.size();
            for (size_t k = 0; k < m; k++) {
                DevSide q{};
                q.p1x = b.hull[k].x; q.p1y = b.hull[k].y;
                q.p2x = b.hull[(k + 1) % m].x; q.p2y = b.hull[(k + 1) % m].y;
                const double L = std::hypot(q.p2x - q.p1x, q.p2y - q.p1y);
                q.nx = (q.p2y - q.p1y) / L;   // counter-clockwise: outward is to the right
                q.ny = -(q.p2x - q.p1x) / L;
                sd.push_back(q);
            }
        }
        for (size_t k = 0; k < sd.size(); k++) {
            DevSide& q = sd[k];
            const double2 a = make_double2(q.p1x, q.p1y), c = make_double2(q.p2x, q.p2y);
            int full = -1, touching = 0;
            for (auto& hc : g->contacts[i]) {
                if (hc.length() == 0.) continue;
                const double2 h1 = make_double2(hc.p1x, hc.p1y), h2 = make_double2(hc.p2x, hc.p2y);
                if (!near_segment(h1, a, c, tol) || !near_segment(h2, a, c, tol)) continue;
                touching++;
                const bool same = (std::hypot(h1.x - a.x, h1.y - a.y) <= tol && std::hypot(h2.x - c.x, h2.y - c.y) <= tol) ||
                                  (std::hypot(h1.x - c.x, h1.y - c.y) <= tol && std::hypot(h2.x - a.x, h2.y - a.y) <= tol);
                if (same) full = hc.box;
            }
            q.nbr = touching == 0 ? SIDE_VOID : (touching == 1 && full >= 0 ? full : SIDE_MIXED);
            q.nid = normal_id(q.nx, q.ny);
            if (k < 4) d.side_nbr[k] = b.hull.empty() ? q.nbr : SIDE_MIXED;
        }
        for (size_t k = sd.size(); k < 4; k++) d.side_nbr[k] = SIDE_MIXED;
        g->srange.push_back(make_int2((int)g->sides.size(), (int)sd.size()));
        g->sides.insert(g->sides.end(), sd.begin(), sd.end());
    }
    g->ptrans.assign(nbox, make_double2(0., 0.));
    Ctx* c = g->ctx;
    HIPCHECK(hipSetDevice(c->device));
    g->d_boxes.upload(g->dboxes, c->stream);
    g->d_contacts.upload(g->dcontacts.empty() ? std::vector<DevContact>(1) : g->dcontacts, c->stream);
    g->d_cseg.upload(g->cseg.empty() ? std::vector<double4>(1) : g->cseg, c->stream);
    g->d_ptrans.upload(g->ptrans, c->stream);
    g->d_sides.upload(g->sides, c->stream);
    g->d_srange.upload(g->srange, c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
}

static void check_kind(int k) {
    if (k < 0 || k > 2) throw Error(BTE_EINVAL, "bad box kind");
}

int bte_geom_create_rects(bte_ctx* ctx, int nbox, const double* rect, const int32_t* kind,
                          const double* Teq, const double* Tinit, bte_geom** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !out || nbox <= 0 || !rect || !kind || !Teq) throw Error(BTE_EINVAL, "bad arguments");
    std::unique_ptr<Geom> g(new Geom());
    g->ctx = c;
    g->boxes.resize(nbox);
    for (int i = 0; i < nbox; i++) {
        RectBox& b = g->boxes[i];
        b.xmin = rect[4 * i]; b.xmax = rect[4 * i + 1]; b.ymin = rect[4 * i + 2]; b.ymax = rect[4 * i + 3];
        if (!(b.xmax > b.xmin && b.ymax > b.ymin)) throw Error(BTE_EGEOM, "degenerate box");
        check_kind(kind[i]);
        b.kind = kind[i];
        b.Teq = Teq[i];
        b.Tinit = Tinit ? Tinit[i] : Teq[i];
    }
    g->contacts.assign(nbox, {});
    rect_contacts(g.get());
    finish_geom(g.get());
    *out = reinterpret_cast<bte_geom*>(g.release());
    BTE_GUARD_END
}

int bte_geom_create_polys(bte_ctx* ctx, int nbox, const int32_t* nvert, const double* xy, const int32_t* kind,
                          const double* Teq, const double* Tinit, bte_geom** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !out || nbox <= 0 || !nvert || !xy || !kind || !Teq) throw Error(BTE_EINVAL, "bad arguments");
    std::unique_ptr<Geom> g(new Geom());
    g->ctx = c;
    g->poly = true;
    g->boxes.resize(nbox);
    size_t off = 0;
    for (int i = 0; i < nbox; i++) {
        RectBox& b = g->boxes[i];
        if (nvert[i] < 3) throw Error(BTE_EGEOM, "box " + std::to_string(i) + ": fewer than 3 vertices");
        std::vector<double2> v(nvert[i]);
        for (int k = 0; k < nvert[i]; k++) {
            v[k] = make_double2(xy[2 * (off + k)], xy[2 * (off + k) + 1]);
            if (!std::isfinite(v[k].x) || !std::isfinite(v[k].y)) throw Error(BTE_EINVAL, "vertices must be finite");
        }
        off += nvert[i];
        b.hull = convex_hull(v);
        if (b.hull.size() < 3) throw Error(BTE_EGEOM, "box " + std::to_string(i) + ": degenerate polygon");
        if (b.hull.size() > POLY_MAXSIDES)
            throw Error(BTE_EGEOM, "box " + std::to_string(i) + ": more than " + std::to_string(POLY_MAXSIDES) + " sides");
        // a vertex on a side between two hull corners makes the reference
        // keep only part of that side as a border (get_sides keeps the
        // first vertex pair of each normal): refused
        for (auto& p : v) {
            bool corner = false;
            for (auto& h : b.hull) corner = corner || (p.x == h.x && p.y == h.y);
            if (corner) continue;
            for (size_t k = 0; k < b.hull.size(); k++)
                if (near_segment(p, b.hull[k], b.hull[(k + 1) % b.hull.size()], 0.))
                    throw Error(BTE_EGEOM, "box " + std::to_string(i) + ": a vertex lies on a side");
        }
        double A = 0., cx = 0., cy = 0.;
        b.xmin = b.ymin = 1e300;
        b.xmax = b.ymax = -1e300;
        for (size_t k = 0; k < b.hull.size(); k++) {
            const double2 p = b.hull[k], q = b.hull[(k + 1) % b.hull.size()];
            const double w = cross2(p.x, p.y, q.x, q.y);
            A += w;
            cx += (p.x + q.x) * w;
            cy += (p.y + q.y) * w;
            b.xmin = std::min(b.xmin, p.x); b.xmax = std::max(b.xmax, p.x);
            b.ymin = std::min(b.ymin, p.y); b.ymax = std::max(b.ymax, p.y);
        }
        b.parea = 0.5 * A;
        b.pcx = cx / (3. * A);
        b.pcy = cy / (3. * A);
        check_kind(kind[i]);
        b.kind = kind[i];
        b.Teq = Teq[i];
        b.Tinit = Tinit ? Tinit[i] : Teq[i];
    }
    double scale = 0.;
    for (auto& b : g->boxes) scale = std::max({scale, std::fabs(b.xmin), std::fabs(b.xmax), std::fabs(b.ymin), std::fabs(b.ymax)});
    const double tol = 1.0e-9 * std::max(scale, 1.0);
    g->contacts.assign(nbox, {});
    for (int i = 0; i < nbox; i++)
        for (int j = 0; j < nbox; j++) {
            if (j == i) continue;
            const RectBox &a = g->boxes[i], &e = g->boxes[j];
            if (a.xmax < e.xmin - tol || e.xmax < a.xmin - tol || a.ymax < e.ymin - tol || e.ymax < a.ymin - tol) continue;
            HostContact hc{j, 0, 0, 0, 0, 0, 0};
            if (poly_contact(a, e, tol, hc)) g->contacts[i].push_back(hc);
        }
    finish_geom(g.get());
    *out = reinterpret_cast<bte_geom*>(g.release());
    BTE_GUARD_END
}

int bte_geom_set_periodic(bte_geom* gg, int box, int target, double tx, double ty) {
    BTE_GUARD_BEGIN
    auto* g = reinterpret_cast<Geom*>(gg);
    if (!g || box < 0 || box >= (int)g->boxes.size()) throw Error(BTE_EINVAL, "bad box id");
    if (g->boxes[box].kind != 2) throw Error(BTE_EINVAL, "only periodic boxes (kind 2) take a translation");
    // translate() (include/geometry_2d.hpp:379-386)
    if (target < 0 || target >= (int)g->boxes.size() || almost_equal(std::hypot(tx, ty), 0.))
        throw Error(BTE_EGEOM, "Error: Bad id for mapping of periodic cells");
    g->dboxes[box].tgt = target;
    g->ptrans[box] = make_double2(tx, ty);
    HIPCHECK(hipSetDevice(g->ctx->device));
    g->d_boxes.upload(g->dboxes, g->ctx->stream);
    g->d_ptrans.upload(g->ptrans, g->ctx->stream);
    HIPCHECK(hipStreamSynchronize(g->ctx->stream));
    BTE_GUARD_END
}

int bte_geom_destroy(bte_geom* gg) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<Geom*>(gg);
    BTE_GUARD_END
}

int bte_geom_contacts(const bte_geom* gg, int maxc, int32_t* ncont, int32_t* nbr, double* geo) {
    BTE_GUARD_BEGIN
    auto* g = reinterpret_cast<const Geom*>(gg);
    if (!g || maxc <= 0 || !ncont || !nbr || !geo) throw Error(BTE_EINVAL, "bad arguments");
    for (size_t i = 0; i < g->boxes.size(); i++) {
        int k = 0;
        for (auto& hc : g->contacts[i]) {
            if (k >= maxc) break;
            nbr[i * maxc + k] = hc.box;
            double* q = geo + (i * maxc + k) * 6;
            q[0] = hc.npx; q[1] = hc.npy; q[2] = hc.p1x; q[3] = hc.p1y; q[4] = hc.p2x; q[5] = hc.p2y;
            k++;
        }
        ncont[i] = k;
    }
    BTE_GUARD_END
}

int bte_krylov_size(int N, const double* B, double dt, double eps, int* m) {
    BTE_GUARD_BEGIN
    if (N < 2 || !B || !m || !(eps > 0)) throw Error(BTE_EINVAL, "bad arguments to bte_krylov_size");
    long double s = 0;
    for (size_t i = 0; i < (size_t)N * N; i++) {
        const long double v = (long double)dt * B[i];
        s += v * v;
    }
    const long double nrm = std::sqrt(s);
    const long double NN = (long double)N * N;
    int k = 2;
    for (; k < 100000; k++) {
        // log of 2 nrm^k e^nrm / Gamma(k), compared in logs to avoid overflow
        const long double lg = std::log(2.0L) + k * std::log(nrm) + nrm - std::lgamma((long double)k);
        if (nrm == 0 || lg - std::log(NN) < std::log((long double)eps)) break;
    }
    *m = k;
    BTE_GUARD_END
}

}  // extern "C"
