// Oracle (test infrastructure only): restatement of the PropagatorBuilder
// Krylov path and of the synthetic collision operator used by the tests.
//   error_bound          include/collision_operator.hpp:216-219
//   Arnoldi_algorithm    src/collision_operator.cpp:815-851 (mode 0 keeps
//                        the reference's j<k projection loop; mode 1 is
//                        textbook MGS over j<=k)
//   build_P              src/collision_operator.cpp:1015-1064
//   energy projection    src/collision_operator.cpp:786-807
//   H.exp()              Eigen MatrixFunctions (external/eigen/unsupported/
//                        Eigen/src/MatrixFunctions/MatrixExponential.h:64-160)
//                        restated as Higham 2005 Pade scaling-and-squaring
#include "orc_common.hpp"
#include <omp.h>

namespace orc {

// --- small dense helpers (column-major) ---
static void matmul(int n, const double* A, const double* B, double* C) {
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) C[i + (size_t)j * n] = 0.;
        for (int k = 0; k < n; k++) {
            double b = B[k + (size_t)j * n];
            if (b == 0.) continue;
            for (int i = 0; i < n; i++) C[i + (size_t)j * n] += A[i + (size_t)k * n] * b;
        }
    }
}

// Solve A X = B (n x n, LU with partial pivoting); B is overwritten by X
static void lu_solve(int n, std::vector<double> A, double* B) {
    std::vector<int> piv(n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = std::fabs(A[k + (size_t)k * n]);
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[i + (size_t)k * n]) > mx) { mx = std::fabs(A[i + (size_t)k * n]); p = i; }
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(A[k + (size_t)j * n], A[p + (size_t)j * n]);
            for (int j = 0; j < n; j++) std::swap(B[k + (size_t)j * n], B[p + (size_t)j * n]);
        }
        double d = A[k + (size_t)k * n];
        if (d == 0.) throw value_error("singular Pade denominator");
        for (int i = k + 1; i < n; i++) {
            double l = A[i + (size_t)k * n] / d;
            A[i + (size_t)k * n] = l;
            if (l == 0.) continue;
            for (int j = k + 1; j < n; j++) A[i + (size_t)j * n] -= l * A[k + (size_t)j * n];
            for (int j = 0; j < n; j++) B[i + (size_t)j * n] -= l * B[k + (size_t)j * n];
        }
    }
    for (int j = 0; j < n; j++)
        for (int i = n - 1; i >= 0; i--) {
            double s = B[i + (size_t)j * n];
            for (int k = i + 1; k < n; k++) s -= A[i + (size_t)k * n] * B[k + (size_t)j * n];
            B[i + (size_t)j * n] = s / A[i + (size_t)i * n];
        }
}

// exp(A) by Pade scaling-and-squaring (Higham, SIAM J. Matrix Anal. Appl.
// 26 (2005) 1179): degree chosen from theta_{3,5,7,9,13} on the 1-norm.
void expm(int n, const double* A, double* E) {
    static const double theta[5] = {1.495585217958292e-2, 2.539398330063230e-1,
                                    9.504178996162932e-1, 2.097847961257068e0,
                                    5.371920351148152e0};
    static const int degs[5] = {3, 5, 7, 9, 13};
    static const double b3[] = {120., 60., 12., 1.};
    static const double b5[] = {30240., 15120., 3360., 420., 30., 1.};
    static const double b7[] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
    static const double b9[] = {17643225600., 8821612800., 2075673600., 302702400., 30270240.,
                                2162160., 110880., 3960., 90., 1.};
    static const double b13[] = {64764752532480000., 32382376266240000., 7771770303897600.,
                                 1187353796428800., 129060195264000., 10559470521600.,
                                 670442572800., 33522128640., 1323241920., 40840800.,
                                 960960., 16380., 182., 1.};
    size_t nn = (size_t)n * n;
    double norm1 = 0;
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += std::fabs(A[i + (size_t)j * n]);
        norm1 = std::max(norm1, s);
    }
    int sel = 4;
    for (int k = 0; k < 4; k++)
        if (norm1 <= theta[k]) { sel = k; break; }
    int s = 0;
    if (sel == 4 && norm1 > theta[4]) s = std::max(0, (int)std::ceil(std::log2(norm1 / theta[4])));
    std::vector<double> X(A, A + nn);
    double scale = std::ldexp(1.0, -s);
    for (auto& x : X) x *= scale;
    std::vector<double> I(nn, 0.);
    for (int i = 0; i < n; i++) I[i + (size_t)i * n] = 1.;
    std::vector<double> A2(nn), A4(nn), A6(nn), U(nn), V(nn), T(nn);
    matmul(n, X.data(), X.data(), A2.data());
    const double* b = nullptr;
    int deg = degs[sel];
    if (deg == 3) b = b3; else if (deg == 5) b = b5; else if (deg == 7) b = b7; else if (deg == 9) b = b9; else b = b13;
    if (deg <= 9) {
        // U = X * sum b_{odd} A^{2k}, V = sum b_{even} A^{2k}
        std::vector<double> Pw(I), Uo(nn, 0.), Ve(nn, 0.);
        for (int k = 0; 2 * k <= deg; k++) {
            for (size_t q = 0; q < nn; q++) {
                Ve[q] += b[2 * k] * Pw[q];
                if (2 * k + 1 <= deg) Uo[q] += b[2 * k + 1] * Pw[q];
            }
            matmul(n, Pw.data(), A2.data(), T.data());
            Pw = T;
        }
        matmul(n, X.data(), Uo.data(), U.data());
        V = Ve;
    } else {
        matmul(n, A2.data(), A2.data(), A4.data());
        matmul(n, A4.data(), A2.data(), A6.data());
        for (size_t q = 0; q < nn; q++) T[q] = b[13] * A6[q] + b[11] * A4[q] + b[9] * A2[q];
        std::vector<double> T2(nn);
        matmul(n, A6.data(), T.data(), T2.data());
        for (size_t q = 0; q < nn; q++)
            T2[q] += b[7] * A6[q] + b[5] * A4[q] + b[3] * A2[q] + b[1] * I[q];
        matmul(n, X.data(), T2.data(), U.data());
        for (size_t q = 0; q < nn; q++) T[q] = b[12] * A6[q] + b[10] * A4[q] + b[8] * A2[q];
        matmul(n, A6.data(), T.data(), V.data());
        for (size_t q = 0; q < nn; q++) V[q] += b[6] * A6[q] + b[4] * A4[q] + b[2] * A2[q] + b[0] * I[q];
    }
    std::vector<double> Num(nn), Den(nn);
    for (size_t q = 0; q < nn; q++) { Num[q] = V[q] + U[q]; Den[q] = V[q] - U[q]; }
    lu_solve(n, Den, Num.data());
    for (int k = 0; k < s; k++) {
        matmul(n, Num.data(), Num.data(), T.data());
        Num = T;
    }
    std::memcpy(E, Num.data(), nn * sizeof(double));
}

// error_bound (collision_operator.hpp:216-219)
long double error_bound(int N, const double* A, int n, double t) {
    long double s = 0;
    for (size_t i = 0; i < (size_t)N * N; i++) s += (long double)(t * A[i]) * (t * A[i]);
    long double nrm = std::sqrt(s);
    return 2.0L * std::pow(nrm, (long double)n) * std::exp(nrm) / std::tgamma((long double)n);
}

int krylov_size(int N, const double* A, double t, double eps) {
    int k = 2;
    while (error_bound(N, A, k, t) / ((long double)N * N) >= eps) {
        k++;
        if (k > 100000) break;
    }
    return k;
}

// Arnoldi_algorithm (collision_operator.cpp:815-851). V: N x n, H: n x n.
void arnoldi(int N, const double* A, const double* b, int n, double t, int mode,
             double* V, double* H) {
    std::fill(H, H + (size_t)n * n, 0.);
    std::fill(V, V + (size_t)N * n, 0.);
    double bn = 0;
    for (int i = 0; i < N; i++) bn += b[i] * b[i];
    bn = std::sqrt(bn);
    std::vector<double> q(N), z(N);
    for (int i = 0; i < N; i++) q[i] = b[i] / bn;
    for (int i = 0; i < N; i++) V[i] = q[i];
    for (int k = 0; k < n; k++) {
        // z = t A q, streaming the column-major A column by column
        std::fill(z.begin(), z.end(), 0.);
        for (int j = 0; j < N; j++) {
            const double qj = q[j];
            const double* a = A + (size_t)j * N;
            for (int i = 0; i < N; i++) z[i] += a[i] * qj;
        }
        for (int i = 0; i < N; i++) z[i] *= t;
        int jmax = mode == 0 ? k : k + 1;
        for (int j = 0; j < jmax; j++) {
            double h = 0;
            for (int i = 0; i < N; i++) h += V[i + (size_t)j * N] * z[i];
            H[j + (size_t)k * n] = h;
            for (int i = 0; i < N; i++) z[i] -= h * V[i + (size_t)j * N];
        }
        double zn = 0;
        for (int i = 0; i < N; i++) zn += z[i] * z[i];
        zn = std::sqrt(zn);
        if (almost_equal(zn, 0., 1.0e-12, 1.0e-9)) break;
        if (k + 1 < n) {
            H[(k + 1) + (size_t)k * n] = zn;
            for (int i = 0; i < N; i++) {
                q[i] = z[i] / zn;
                V[i + (size_t)(k + 1) * N] = q[i];
            }
        }
    }
}

// build_P (collision_operator.cpp:1015-1064): P(:,i) = (V exp(H))(:,0)
// columns [c0, c1) of P into P[:, 0 : c1 - c0] (OpenMP over columns)
void build_P_cols(int N, const double* A, double t, int m, int mode, int c0, int c1, double* P) {
#pragma omp parallel
    {
        std::vector<double> V((size_t)N * m), H((size_t)m * m), E((size_t)m * m), b(N);
#pragma omp for schedule(dynamic, 1)
        for (int i = c0; i < c1; i++) {
            std::fill(b.begin(), b.end(), 0.);
            b[i] = 1.0;
            arnoldi(N, A, b.data(), m, t, mode, V.data(), H.data());
            expm(m, H.data(), E.data());
            for (int r = 0; r < N; r++) {
                double s = 0;
                for (int k = 0; k < m; k++) s += V[r + (size_t)k * N] * E[k];
                P[r + (size_t)(i - c0) * N] = s;
            }
        }
    }
}

void build_P(int N, const double* A, double t, int m, int mode, double* P) {
    build_P_cols(N, A, t, m, mode, 0, N, P);
}

// Column energy projection (collision_operator.cpp:786-807): subtract the
// compensated column mean so each column sums to zero.
void project_columns(int N, double* B) {
    for (int j = 0; j < N; j++) {
        double* c = B + (size_t)j * N;
        double d = -neumaier_sum(c, N) / N;
        for (int i = 0; i < N; i++) c[i] += d;
    }
}

// Synthetic collision operator: B = -Omega, Omega = diag(w) - S diag(C)^-1,
// S symmetric >= 0 with a seeded pattern of the given density, scaled
// symmetrically so that S 1 = C .* w (then 1^T Omega = 0 and Omega C = 0).
// The pattern and magnitudes come from PCG32(seed, stream 0x6f6d).
void synthetic_B(int N, const double* C, const double* w, double density,
                 uint64_t seed, int sinkhorn_iters, double* B) {
    pcg32 rng(seed, 0x6f6du);
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
    for (int it = 0; it < sinkhorn_iters; it++) {
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
        // K_ij = S_ij / C_j; B = K - diag(row sums / C) so columns sum to 0
        B[ri[e] + (size_t)cj[e] * N] = s / C[cj[e]];
        B[cj[e] + (size_t)ri[e] * N] = s / C[ri[e]];
        srow[ri[e]] += s;
        srow[cj[e]] += s;
    }
    // diagonal: -sum_i K_ij so that 1^T B = 0 exactly in exact arithmetic;
    // equals -w_j once the scaling has converged
    for (int j = 0; j < N; j++) B[j + (size_t)j * N] = -srow[j] / C[j];
}

}  // namespace orc
