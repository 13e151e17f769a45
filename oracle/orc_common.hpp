// Oracle: CPU restatement of BTE-Barna's deviational Monte Carlo and Krylov
// propagator algorithms. TEST INFRASTRUCTURE ONLY: it is linked by the tests,
// by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
// nothing that ships. It restates the reference's algorithms in plain C++17
// (no Boost, TBB, Eigen or msgpack, which the image lacks), citing the
// reference file:line each piece follows.
//
// Parity status per function is listed in the header of orc_capi.cpp.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <tuple>

namespace orc {

// include/constants.hpp:37,47,56
constexpr double kB = 1.3806488e-23;
constexpr double hbar = 1.054571726e-34;
constexpr double PI = 3.14159265358979323846;
constexpr double nsigma = 2.828427124746190097603377448419396157138;

struct geometry_error : std::runtime_error {
    explicit geometry_error(const std::string& s) : std::runtime_error(s) {}
};
struct value_error : std::runtime_error {
    explicit value_error(const std::string& s) : std::runtime_error(s) {}
};
struct input_error : std::runtime_error {
    explicit input_error(const std::string& s) : std::runtime_error(s) {}
};

// PCG32 = pcg_engines::setseq_xsh_rr_64_32 (external/pcg-cpp-0.98
// include/pcg_random.hpp:1675 family), seeded like pcg32(state, stream):
// inc = stream<<1|1, state = bump(state + inc), output taken from the
// pre-advance state.
struct pcg32 {
    uint64_t state = 0, inc = 1;
    pcg32() = default;
    pcg32(uint64_t initstate, uint64_t initseq) {
        inc = (initseq << 1u) | 1u;
        state = (initstate + inc) * 6364136223846793005ULL + inc;
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    // pcg_extras::bounded_rand (pcg_extras.hpp:517-528)
    uint32_t bounded(uint32_t bound) {
        uint32_t threshold = (0u - bound) % bound;
        for (;;) {
            uint32_t r = next();
            if (r >= threshold) return r % bound;
        }
    }
    // 53-bit uniform in [0,1) from two draws
    double uniform() {
        uint64_t a = next() >> 5, b = next() >> 6;
        return (double)((a << 26) | b) * (1.0 / 9007199254740992.0);
    }
};

// include/utilities.hpp:214-223
inline bool almost_equal(double a, double b, double abstol = 1e-8,
                         double reltol = 1e-5) {
    const double tol = std::fabs(abstol) + std::fabs(reltol * b);
    return std::fabs(a - b) < tol;
}

// include/utilities.hpp:385-389
inline double bose_einstein(double omega, double T) {
    constexpr double prefactor = 1e12 * hbar / kB;
    return 1. / (std::exp(prefactor * omega / T) - 1.);
}

// include/utilities.hpp:401-415
inline double bose_einstein_kernel(double omega, double T) {
    constexpr double prefactor = 0.5 * 1e12 * hbar / kB;
    double x = prefactor * omega / T;
    if (x < 1e-320) return 1.;
    if (x > 300.) return 0.;
    double s = x / std::sinh(x);
    return s * s;
}

// include/utilities.hpp:59-75 (Neumaier compensated sum)
inline double neumaier_sum(const double* a, size_t n) {
    double sum = 0, c = 0;
    for (size_t i = 0; i < n; i++) {
        double t = sum + a[i];
        if (std::fabs(sum) >= std::fabs(a[i]))
            c += (sum - t) + a[i];
        else
            c += (a[i] - t) + sum;
        sum = t;
    }
    return sum + c;
}

// A flat phonon mode table: the reference's Gamma_grid with nq q-points and
// nb = n/nq bands; mode index = iq*nb + ib (beyondRTA_MC2d.cpp:1120).
struct Modes {
    int n = 0;
    int nq = 1;
    std::vector<double> omega, vx, vy, w;
    double V = 1.0;      // unit-cell volume [nm^3]
    double lattz = 1.0;  // lattvec(2,2) [nm]
};

// get_v: velocities that are almost_equal to 0 are set to 0
// (RTA_MC2D.cpp:843-856; beyondRTA_MC2d.cpp:686-699)
inline void get_v(const Modes& m, int mode, double& vx, double& vy) {
    vx = m.vx[mode];
    vy = m.vy[mode];
    if (almost_equal(vx, 0.)) vx = 0.;
    if (almost_equal(vy, 0.)) vy = 0.;
}

// Grid_distribution (sampling.hpp:50-115, sampling.cpp:26-44): normalised
// cumulative + lower_bound sampling; optional per-entry sign.
struct GridDist {
    std::vector<double> cumulative;
    std::vector<int8_t> signs;
    double cumulsum = -1.0;
    int nq = 1;
    void fill_cumulative(const std::vector<double>& p) {
        for (double x : p)
            if (x < 0.) throw value_error("probabilities must be >= 0.");
        cumulative.resize(p.size());
        double acc = 0;
        for (size_t i = 0; i < p.size(); i++) {
            acc += p[i];
            cumulative[i] = acc;
        }
        double last = cumulative.back();
        for (auto& c : cumulative) c /= last;
    }
    size_t sample(pcg32& rng) const {
        double u = rng.uniform();
        return std::lower_bound(cumulative.begin(), cumulative.end(), u) -
               cumulative.begin();
    }
};

// BE_derivative_distribution (sampling.cpp:47-68): pmf ~ kernel(w,T)*rate
inline GridDist be_derivative(const Modes& m, double T) {
    GridDist d;
    d.nq = m.nq;
    std::vector<double> pmf(m.n);
    for (int i = 0; i < m.n; i++)
        pmf[i] = m.omega[i] <= 0. ? 0. : bose_einstein_kernel(m.omega[i], T) * m.w[i];
    d.fill_cumulative(pmf);
    return d;
}

// Isothermal_wall_distribution (sampling.cpp:189-269); normal points into
// the material the particles are emitted into.
inline GridDist isothermal_wall(const Modes& m, double Twall, double Teq,
                                double nx, double ny) {
    GridDist d;
    d.nq = m.nq;
    std::vector<double> pmf(m.n);
    double acc = 0.;
    for (int i = 0; i < m.n; i++) {
        double vn = nx * m.vx[i] + ny * m.vy[i];
        if (vn <= 0. || almost_equal(0., m.omega[i]))
            pmf[i] = 0.;
        else
            pmf[i] = vn * m.omega[i] *
                     std::fabs(bose_einstein(m.omega[i], Twall) -
                               bose_einstein(m.omega[i], Teq));
        acc += pmf[i];
    }
    d.cumulsum = 1.0e12 * hbar * acc;
    d.signs.assign(m.n, (Twall - Teq) >= 0 ? 1 : -1);
    d.fill_cumulative(pmf);
    return d;
}

// Isothermal_wall_distribution::get_flux (sampling.cpp:266-272) [J/(nm^2 ps)]
inline double wall_flux(const GridDist& d, double vuc) {
    return d.signs[0] * d.cumulsum / (d.nq * vuc);
}

// Ntogenerate with Bernoulli remainder (sampling.cpp:248-262, 313-325)
inline size_t n_to_generate(double gen, pcg32& rng) {
    size_t n = (size_t)gen;
    double rem = gen - std::floor(gen);
    if (rng.uniform() <= rem) n++;
    return n;
}

// outTref_distribution (sampling.cpp:271-330)
// ref_distribution (sampling.cpp:70-93): a given signed distribution fd over
// the modes after the ngamma Gamma entries (mode ngamma + j <- fd[j]);
// pmf |fd|, sign of fd, cumulsum = sum |fd|
inline GridDist ref_distribution(const Modes& m, const std::vector<double>& fd, int ngamma) {
    if ((int)fd.size() + ngamma != m.n) throw value_error("given distribution of the wrong length");
    GridDist d;
    d.nq = m.nq;
    std::vector<double> pmf(m.n, 0.);
    d.signs.assign(m.n, 1);
    double cs = 0;
    for (size_t j = 0; j < fd.size(); j++) {
        const int k = (int)j + ngamma;
        d.signs[k] = std::signbit(fd[j]) ? -1 : 1;
        pmf[k] = std::fabs(fd[j]);
        cs += std::fabs(fd[j]);
    }
    d.cumulsum = cs;
    d.fill_cumulative(pmf);
    return d;
}

inline GridDist out_tref(const Modes& m, double Tinit, double Tref) {
    if (Tinit <= 0 || Tref <= 0) throw value_error("Invalid temperature value");
    GridDist d;
    d.nq = m.nq;
    std::vector<double> pmf(m.n, 0.);
    double acc = 0;
    for (int i = 0; i < m.n; i++) {
        if (almost_equal(m.omega[i], 0.)) continue;
        pmf[i] = m.omega[i] * std::fabs(bose_einstein(m.omega[i], Tinit) -
                                        bose_einstein(m.omega[i], Tref));
        acc += pmf[i];
    }
    d.cumulsum = 1.0e12 * hbar * acc;
    d.signs.assign(m.n, (Tinit - Tref) >= 0 ? 1 : -1);
    d.fill_cumulative(pmf);
    return d;
}

// calc_cv (bulk_properties.hpp:32-47) [J/(K nm^3)]
inline double calc_cv(const Modes& m, double T) {
    double acc = 0;
    for (int i = 0; i < m.n; i++) acc += bose_einstein_kernel(m.omega[i], T);
    return acc * kB / m.nq / m.V;
}

// Natural cubic spline (src/interpolation.cpp:29-100)
struct CubicSpline1D {
    std::vector<double> a, b, c, d, x0, xe;
    CubicSpline1D() = default;
    CubicSpline1D(const std::vector<double>& X, const std::vector<double>& Y) {
        int n = (int)Y.size() - 1;
        std::vector<double> A(Y), B(n), D(n), h(n), alpha(n, 0.), C(n + 1),
            l(n + 1), mu(n + 1), z(n + 1);
        for (int i = 0; i < n; i++) h[i] = X[i + 1] - X[i];
        for (int i = 1; i < n; i++)
            alpha[i] = 3. / h[i] * (A[i + 1] - A[i]) - 3. / h[i - 1] * (A[i] - A[i - 1]);
        l[0] = 1;
        mu[0] = 0;
        z[0] = 0;
        for (int i = 1; i < n; i++) {
            l[i] = 2 * (X[i + 1] - X[i - 1]) - h[i - 1] * mu[i - 1];
            mu[i] = h[i] / l[i];
            z[i] = (alpha[i] - h[i - 1] * z[i - 1]) / l[i];
        }
        l[n] = 1;
        z[n] = 0;
        C[n] = 0;
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
    double interpolate(double nx) const {
        size_t idx = std::lower_bound(xe.begin(), xe.end(), nx) - xe.begin();
        if (idx >= a.size()) idx--;
        double dx = nx - x0[idx];
        return a[idx] + dx * (b[idx] + dx * (c[idx] + d[idx] * dx));
    }
};

}  // namespace orc
