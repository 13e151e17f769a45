// Host-side runtime objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>
#include <map>
#include <memory>
#include "../../include/bte_abi.h"
#include "device_common.hpp"

namespace bte {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};

void set_error(const std::string& s);

#define HIPCHECK(x)                                                                 \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess)                                                       \
            throw ::bte::Error(BTE_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BTE_GUARD_BEGIN try {
#define BTE_GUARD_END                                 \
    }                                                 \
    catch (const ::bte::Error& e) {                   \
        ::bte::set_error(e.what());                   \
        return e.code;                                \
    }                                                 \
    catch (const std::exception& e) {                 \
        ::bte::set_error(e.what());                   \
        return BTE_EINVAL;                            \
    }                                                 \
    return BTE_OK;

// Device buffer owned by a handle.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(BTE_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        HIPCHECK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0;
    // BLAS handle kept for the context's lifetime (created on first use)
    void* blas = nullptr;
    void (*blas_free)(void*) = nullptr;
};

struct Modes {
    Ctx* ctx = nullptr;
    int n = 0, nq = 1;
    std::vector<double> omega, vx, vy, w;
    std::vector<double> sigma;  // per-mode broadening (base_sigma), empty when not given
    double V = 1.0, lattz = 1.0;
    DevBuf<double> d_vx, d_vy, d_w, d_omega;
};

// A box: an axis-aligned rectangle, or a convex polygon (hull not empty:
// its vertices counter-clockwise, xmin..ymax its bounding box)
struct RectBox {
    double xmin, xmax, ymin, ymax;
    int kind;
    double Teq, Tinit;
    std::vector<double2> hull;
    double parea = 0., pcx = 0., pcy = 0.;  // polygon area and centroid
    double cx() const { return hull.empty() ? 0.5 * (xmin + xmax) : pcx; }
    double cy() const { return hull.empty() ? 0.5 * (ymin + ymax) : pcy; }
    double area() const { return hull.empty() ? (xmax - xmin) * (ymax - ymin) : parea; }
};
struct HostContact {
    int box;
    double npx, npy, p1x, p1y, p2x, p2y;
    double length() const { return std::hypot(p2x - p1x, p2y - p1y); }
};

struct Geom {
    Ctx* ctx = nullptr;
    std::vector<RectBox> boxes;
    std::vector<std::vector<HostContact>> contacts;  // ordered by neighbour id
    std::vector<DevBox> dboxes;
    std::vector<DevContact> dcontacts;
    std::vector<double4> cseg;     // per contact: shared segment extended by 1e-6
    std::vector<double2> ptrans;   // per box: periodic translation (0 otherwise)
    // sides of every box (rectangles: left, bottom, top, right at 4 i ..
    // 4 i + 3), their ranges per box and the distinct outward normals
    bool poly = false;             // some box is not an axis-aligned rectangle
    std::vector<DevSide> sides;
    std::vector<int2> srange;
    std::vector<double2> normals;
    DevBuf<DevBox> d_boxes;
    DevBuf<DevContact> d_contacts;
    DevBuf<double4> d_cseg;
    DevBuf<double2> d_ptrans;
    DevBuf<DevSide> d_sides;
    DevBuf<int2> d_srange;
};

// Propagators of one material at ascending temperatures; a single entry with
// T <= 0 is used at every box temperature.
struct Prop {
    Ctx* ctx = nullptr;
    int ncol = 0, ngamma = 0;
    std::vector<double> T;
    std::vector<std::vector<double>> P;  // host copies (column-major), per T
    DevBuf<uint2> d_alias;  // per T: ncol x ncol entries, column-major
};

// rocBLAS handle of a context (krylov.hip), created on first use and bound
// to the context's stream
void* blas_handle(Ctx* c);

// ---- host helpers (host_util.cpp) ----
double h_bose_einstein(double omega, double T);
double h_be_kernel(double omega, double T);
double h_calc_cv(const Modes& m, double T);
// Vose alias construction; pmf >= 0, returns false if the pmf sums to 0
bool build_alias(const double* pmf, int n, uint32_t* thresh, uint32_t* alias);
// The distribution of propagator_H::get(R) (collision_operator.cpp:892-904)
// for R uniform in [0, 1) over a signed cumulative sc[n]: pick k with
// probability pmf[k]; a lower_bound past the end picks n-1.
void pick_pmf(const double* sc, int n, double* pmf);
// Spectral weight matrix W [n][3 nomega + nq] (row-major) on the frequency
// ticks grid: Gamma densities gamma_pdf(omega^2/sigma, sigma/omega, tick)
// of every mode (RTA_MC2D.cpp:61-63, 723-731; dist_reader.cpp:349-366), the
// same times vx and vy, and 1/omega in the mode's q column. sigma per mode,
// or sigma_default for all, raised to exp(p25 - 1.5 (p75 - p25)) of log
// sigma; modes with omega ~ 0 get no weights.
void spectral_weights(const Modes& m, const double* sigma, double sigma_default, int nomega, const double* grid,
                      std::vector<double>& W);
// calculate_gradientT (geometry_2d.cpp:524-576) per box [K/nm]
std::vector<double2> gradient_T(const struct Geom& g);
// isothermal wall pmf (sampling.cpp:189-269) into pmf; returns cumulsum
double wall_pmf(const Modes& m, double Twall, double Teq, double nx, double ny, double* pmf);
// natural cubic spline coefficients (src/interpolation.cpp:29-100):
// per segment (a, b, c, d, x0) and segment right ends xe
struct Spline {
    std::vector<double> a, b, c, d, x0, xe;
    void build(const std::vector<double>& X, const std::vector<double>& Y);
    double eval(double x) const;
};

}  // namespace bte
