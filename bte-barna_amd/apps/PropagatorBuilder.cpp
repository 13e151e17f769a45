// PropagatorBuilder front-end on the GPU engine (src/PropagatorBuilder.cpp
// 56-170): P = exp(dt B) from an energy-conserving operator B.
//
// usage: PropagatorBuilder <B.eigen.bin | synthetic:N:seed[:density]> <T> <dt>
//                          [nthreads] [m] [fp64|fp32|bf16|pade] [reference|textbook]
//
// The reference assembles B from an HDF5 database and third-order force
// constants; the GPU engine starts from the operator itself: a saved
// <stem>.B.eigen.bin, or the seeded synthetic operator of an N-mode
// synthetic material at T. Writes <stem>_<T>K_<dt>ps.P.eigen.bin (and the
// .B.eigen.bin for a synthetic operator). m defaults to the reference's
// Krylov size rule (error_bound per element below 1e-8). "pade" replaces
// the Krylov build by the FP64 Pade(13) scaling-and-squaring exponential of
// the whole matrix (exact to rounding; about half the FP64 Krylov time for
// all columns of a 8192-mode operator at m = 30).
#include "front_common.hpp"
#include <algorithm>

using namespace front;

static int run(int argc, char** argv) {
    if (argc < 4 || argc > 8) {
        std::cerr << "Usage: " << argv[0]
                  << " <B.eigen.bin|synthetic:N:seed[:density]> <T> <dt> [nthreads] [m] [fp64|fp32|bf16|pade]"
                     " [reference|textbook]"
                  << std::endl;
        return 1;
    }
    const std::string src = argv[1];
    const double T = std::atof(argv[2]), dt = std::atof(argv[3]);
    int m = argc > 5 ? std::atoi(argv[5]) : 0;
    const std::string pr = argc > 6 ? argv[6] : "fp64";
    const std::string am = argc > 7 ? argv[7] : "reference";
    const int prec = pr == "fp32" ? BTE_PREC_FP32 : (pr == "bf16" ? BTE_PREC_BF16 : BTE_PREC_FP64);
    if (pr != "fp64" && pr != "fp32" && pr != "bf16" && pr != "pade")
        throw std::runtime_error("precision must be fp64, fp32, bf16 or pade");
    const bool pade = pr == "pade";
    if (am != "reference" && am != "textbook") throw std::runtime_error("Arnoldi mode must be reference or textbook");
    if (!(T > 0) || !(dt > 0)) throw std::runtime_error("T and dt must be positive");
    std::vector<double> B;
    int64_t N = 0;
    std::string stem;
    if (src.rfind("synthetic:", 0) == 0) {
        int n = 0;
        unsigned long long seed = 1;
        double density = 0.05;
        if (std::sscanf(src.c_str(), "synthetic:%d:%llu:%lf", &n, &seed, &density) < 2 || n < 2)
            throw std::runtime_error("synthetic source is synthetic:N:seed[:density]");
        std::vector<double> om(n), vx(n), vy(n), w(n), C(n);
        check(bte_synthetic_modes(n, seed, 0.1, 15.0, 0.5, 8.0, 1.0, 1000.0, 0, om.data(), vx.data(), vy.data(),
                                  w.data()),
              "synthetic modes");
        for (int i = 0; i < n; i++) C[i] = kB * be_kernel(om[i], T);
        N = n;
        B.resize((size_t)N * N);
        check(bte_synthetic_B(n, C.data(), w.data(), density, seed, 200, B.data()), "synthetic operator");
        stem = "synthetic_" + std::to_string(n) + "_" + std::to_string(seed);
    } else {
        B = load_eigen_bin(src, N);
        stem = src;
        const std::string suf = ".B.eigen.bin";
        if (stem.size() > suf.size() && stem.compare(stem.size() - suf.size(), suf.size(), suf) == 0)
            stem = stem.substr(0, stem.size() - suf.size());
    }
    if (m <= 0 && !pade) {
        check(bte_krylov_size((int)N, B.data(), dt, 1e-8, &m), "Krylov size");
        std::cout << "# Krylov subspace size " << m << std::endl;
        if (m > 64) throw std::runtime_error("Krylov size above 64: pass m explicitly");
    }
    bte_ctx* ctx = nullptr;
    check(bte_ctx_create(0, &ctx), "GPU context");
    std::vector<double> P((size_t)N * N);
    bte_krylov_report rep{};
    if (pade) {
        check(bte_expm(ctx, (int)N, B.data(), dt, P.data(), &rep.build_ms), "Pade exponential");
        rep.gemm_tflops = 0.;
        for (int64_t j = 0; j < N; j++) {
            double sum = 0.;
            for (int64_t i = 0; i < N; i++) sum += P[(size_t)j * N + i];
            rep.max_col_sum_err = std::max(rep.max_col_sum_err, std::fabs(sum - 1.));
        }
    } else {
        check(bte_krylov_build(ctx, (int)N, B.data(), dt, m, prec,
                               am == "textbook" ? BTE_ARNOLDI_TEXTBOOK : BTE_ARNOLDI_REFERENCE, P.data(), &rep),
              "Krylov build");
    }
    char tag[128];
    std::snprintf(tag, sizeof tag, "_%gK_%gps", T, dt);
    if (src.rfind("synthetic:", 0) == 0) save_eigen_bin(stem + tag + ".B.eigen.bin", N, B.data());
    save_eigen_bin(stem + tag + ".P.eigen.bin", N, P.data());
    std::cout << "# P written to " << stem + tag + ".P.eigen.bin" << ": N = " << N << ", " << (pade ? std::string("FP64 Pade(13)") : "m = " + std::to_string(m) + ", " + pr)
              << ", " << rep.build_ms << " ms, GEMM " << rep.gemm_tflops << " TFLOP/s, max |colsum - 1| "
              << rep.max_col_sum_err << std::endl;
    bte_ctx_destroy(ctx);
    return 0;
}

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "PropagatorBuilder: " << e.what() << std::endl;
        return 1;
    }
}
