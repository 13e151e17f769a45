// Krylov propagator P = exp(dt*B) for all columns at once.
//
// Reference: build_P (src/collision_operator.cpp:1015-1064) runs one
// Arnoldi_algorithm (815-851) per column with b = e_i and keeps
// P(:,i) = (V exp(H))(:,0). Here every column advances together: step k is
// one N x N x N GEMM Z = dt*B*V_k (FP64 and FP32 on rocBLAS; BF16 hi+lo
// split as one hand-written MFMA kernel with FP32 accumulation, A_hi V_hi +
// A_hi V_lo + A_lo V_hi), followed by a per-column FP64 modified
// Gram-Schmidt kernel, a per-column FP64 exponential of the small Hessenberg
// matrix and a combine kernel that forms P(:,i) and applies the FP64 column
// energy projection (collision_operator.cpp:786-807 applied to P: columns
// sum to one).
#include "runtime.hpp"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <hip/hip_bf16.h>
#include <chrono>
#include <cstring>

namespace bte {

#define RBCHECK(x)                                                                    \
    do {                                                                              \
        rocblas_status s_ = (x);                                                      \
        if (s_ != rocblas_status_success)                                             \
            throw ::bte::Error(BTE_EHIP, std::string(#x) + ": rocblas status " +       \
                                             std::to_string((int)s_));                \
    } while (0)

// FP64 -> compute format
__global__ void kr_to_f32(const double* __restrict__ a, float* __restrict__ o, size_t n, double scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = (float)(a[i] * scale);
}
// column-major N x N a -> Np x Np FP32, column stride ldo (rows and columns
// past N are not written)
__global__ void kr_to_f32_ld(const double* __restrict__ a, float* __restrict__ o, int N, size_t ldo, double scale) {
    const size_t n = (size_t)N * N;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        o[(q / N) * ldo + q % N] = (float)(a[q] * scale);
}
// column-major N x N a -> the row-major hi/lo-interleaved operand of
// kr_bf16x3_gemm: row r holds, for every 8 consecutive k, the 8 bf16 hi
// parts of (float)(scale * a[r + k N]) then the 8 lo parts; zero past N up to
// Np. Rows ld bf16 apart. 32 x 32 tiles transposed through LDS (grid Np/32 x
// Np/32, 256 threads).
__global__ __launch_bounds__(256) void kr_to_bf16_il(const double* __restrict__ a, __hip_bfloat16* __restrict__ A,
                                                     int N, size_t ld, double scale) {
    __shared__ float t[32][33];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int kk = ty; kk < 32; kk += 8) {
        const int r = r0 + tx, k = k0 + kk;
        t[kk][tx] = (r < N && k < N) ? (float)(a[r + (size_t)k * N] * scale) : 0.f;
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {
        const int r = r0 + rr, k = k0 + tx;
        const float v = t[tx][rr];
        const __hip_bfloat16 h = __float2bfloat16(v);
        __hip_bfloat16* e = A + (size_t)r * ld + 2 * (k & ~7) + (k & 7);
        e[0] = h;
        e[8] = __float2bfloat16(v - __bfloat162float(h));
    }
}

// Z = A_hi V_hi + A_hi V_lo + A_lo V_hi in FP32 on MFMA: the compensated BF16
// form of the block product dt*B*V_k that replaces Arnoldi_algorithm's
// per-column t*A*q (collision_operator.cpp:832). A (Np x Np, rows ld_a bf16
// apart) and V_k (Np columns, ld_v bf16 apart) are hi/lo interleaved per 8
// consecutive k, so one 32-deep stage of a row is one 128-byte line holding
// both parts. Z: column-major Np x Np. Np a multiple of 256.
// Workgroup tile 256 x 256, 8 waves of 128 x 64 (4 x 2 blocks of
// v_mfma_f32_32x32x16_bf16, three products each per 16-deep step); stages of
// K = 32 (64 KB) through a 2-stage LDS ring filled by 16-byte direct-to-LDS
// loads, issued for stage s+1 while stage s is multiplied. In LDS a row is 8
// chunks of 16 bytes, chunk c stored at c ^ ((row >> 1) & 7): the 16 lanes
// of a 128-bit fragment read then cover all 64 banks.
namespace gk {
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;
constexpr int WM = 2, WN = 4, MB = 4, NB = 2;       // waves and blocks per wave
constexpr int TM = WM * MB * 32, TN = WN * NB * 32;  // 256 x 256
constexpr int NW = WM * WN, NT = 64 * NW;            // 8 waves
constexpr int ROWB = 128, STAGE = (TM + TN) * ROWB;  // 64 KB
constexpr int LDS = 2 * STAGE;
constexpr int LPW = STAGE / 1024 / NW;               // 1 KB direct loads per wave per stage
__device__ inline int swz(int r) { return (r >> 1) & 7; }
}  // namespace gk

__global__ __launch_bounds__(gk::NT) void kr_bf16x3_gemm(const __hip_bfloat16* __restrict__ A, size_t lda,
                                                        const __hip_bfloat16* __restrict__ V, size_t ldv,
                                                        float* __restrict__ Z, int Np) {
    using namespace gk;
    extern __shared__ __align__(16) unsigned char lds[];
    const int tmn = Np / TM, tnn = Np / TN, total = tmn * tnn;
    // workgroups go round-robin to the 8 XCDs (workgroup b to XCD b % 8).
    // From 16 x 16 tiles up, XCD x owns a quarter of the tile rows times half
    // of the tile columns and walks it in groups of 4 row tiles x 8 column
    // tiles: the 32 tiles it runs at once share 4 row panels and 8 column
    // panels in its L2, and over the whole product it reads a quarter of A and
    // half of V (a band of whole tile rows per XCD would read all of V).
    // Smaller grids: contiguous eighths in bands of 4 row tiles, or row order.
    const int b = blockIdx.x;
    const int u = (total % 8 == 0) ? (b % 8) * (total / 8) + b / 8 : b;
    int tr, tc;
    if (tmn % 16 == 0 && tnn % 16 == 0) {
        const int x = b % 8, t = b / 8;
        const int bm = tmn / 4, bn = tnn / 2;  // the XCD's block of tiles
        const int gpr = bn / 8;                 // groups of 8 columns per band of 4 rows
        const int grp = t / 32, w = t % 32;
        tr = (x / 2) * bm + (grp / gpr) * 4 + w % 4;
        tc = (x % 2) * bn + (grp % gpr) * 8 + w / 4;
    } else if (tmn % 4 == 0) {
        const int band = u / (4 * tnn), w = u % (4 * tnn);
        tr = band * 4 + w % 4;
        tc = w / 4;
    } else {
        tr = u / tnn;
        tc = u % tnn;
    }
    const int row0 = tr * TM, col0 = tc * TN;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // direct load q of this wave fills 1 KB block blk of a stage: 8 rows,
    // lane i writes 16 bytes at 16 i: row i / 8, stored chunk i % 8, which
    // holds chunk (i % 8) ^ swz(row) of the row's 128 bytes
    const __hip_bfloat16* gsrc[LPW];
    int ldsoff[LPW];
#pragma unroll
    for (int q = 0; q < LPW; q++) {
        const int blk = q * NW + wv;
        const __hip_bfloat16* base;
        int r;
        if (blk < TM / 8) {
            r = blk * 8 + (lane >> 3);
            base = A + (size_t)(row0 + r) * lda;
        } else {
            r = (blk - TM / 8) * 8 + (lane >> 3);
            base = V + (size_t)(col0 + r) * ldv;
        }
        gsrc[q] = base + ((lane & 7) ^ swz(r)) * 8;
        ldsoff[q] = blk * 1024;
    }
    auto issue = [&](int s, int buf) {
#pragma unroll
        for (int q = 0; q < LPW; q++)
            __builtin_amdgcn_global_load_lds((const void*)(gsrc[q] + (size_t)s * 64),
                                             (lds_void*)(lds + buf * STAGE + ldsoff[q]), 16, 0, 0);
    };
    f32x16 acc[MB][NB];
#pragma unroll
    for (int x = 0; x < MB; x++)
#pragma unroll
        for (int y = 0; y < NB; y++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[x][y][e] = 0.f;
    const int wm = wv % WM, wn = wv / WM;
    const int fr = lane & 31, fh = lane >> 5;
    // fragment of 16-deep step ks for this lane: k chunk j = 2 ks + fh of
    // row fr, hi at chunk 2 j and lo at chunk 2 j + 1
    int ahoff[MB][2], aloff[MB][2], bhoff[NB][2], bloff[NB][2];
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
        const int r = wm * MB * 32 + mb * 32 + fr;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int j = 2 * ks + fh;
            ahoff[mb][ks] = r * ROWB + (((2 * j) ^ swz(r)) * 16);
            aloff[mb][ks] = r * ROWB + (((2 * j + 1) ^ swz(r)) * 16);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int r = wn * NB * 32 + nb * 32 + fr;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int j = 2 * ks + fh;
            bhoff[nb][ks] = TM * ROWB + r * ROWB + (((2 * j) ^ swz(r)) * 16);
            bloff[nb][ks] = TM * ROWB + r * ROWB + (((2 * j + 1) ^ swz(r)) * 16);
        }
    }
    const int nk = Np / 32;
    issue(0, 0);
    for (int s = 0; s < nk; s++) {
        const int buf = s & 1;
        // every wave's loads of stage s have landed (the release fence of
        // __syncthreads waits for this wave's outstanding loads, which with
        // two buffers are exactly stage s), and every wave is done with stage
        // s-1, whose buffer then takes stage s+1; the fences also keep the
        // compiler from moving LDS reads across the barrier
        __syncthreads();
        if (s + 1 < nk) issue(s + 1, buf ^ 1);
        const unsigned char* sb = lds + buf * STAGE;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x8 ah[MB], al[MB], bh[NB], bl[NB];
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                ah[mb] = *reinterpret_cast<const bf16x8*>(sb + ahoff[mb][ks]);
                al[mb] = *reinterpret_cast<const bf16x8*>(sb + aloff[mb][ks]);
            }
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                bh[nb] = *reinterpret_cast<const bf16x8*>(sb + bhoff[nb][ks]);
                bl[nb] = *reinterpret_cast<const bf16x8*>(sb + bloff[nb][ks]);
            }
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bh[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bl[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mb], bh[nb], acc[mb][nb], 0, 0, 0);
        }
    }
    // a 32 x 32 block: lane holds column lane & 31, rows 8 q + 4 (lane >> 5) + e
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int col = col0 + wn * NB * 32 + nb * 32 + fr;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = row0 + wm * MB * 32 + mb * 32 + 8 * q + 4 * fh;
                *reinterpret_cast<float4*>(Z + (size_t)col * Np + row) = make_float4(
                    acc[mb][nb][4 * q], acc[mb][nb][4 * q + 1], acc[mb][nb][4 * q + 2], acc[mb][nb][4 * q + 3]);
            }
        }
}
__global__ void kr_scale64(const double* __restrict__ a, double* __restrict__ o, size_t n, double scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = a[i] * scale;
}
__global__ void kr_f32_to_f64(const float* __restrict__ a, double* __restrict__ o, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = (double)a[i];
}

// Basis storage: FP64 (prec 0), FP32 (1) or BF16 hi/lo (2). Element r of
// column i of V_j has the logical index (j*nc + i)*cs + r (nc columns per
// step, cs a multiple of 8). FP64/FP32: stored at that index (nc = cs = N).
// BF16: every 8 consecutive elements are stored as 8 hi then 8 lo values,
// so logical index x lives at hi[il(x)] and lo[il(x)] = hi[il(x) + 8] with
// il(x) = 2 (x & ~7) + (x & 7); nc = Np (N rounded up to the product
// kernel's tile), cs = Np + 32 (columns 2 Np + 64 bf16 apart in memory).
struct Basis {
    double* f64;
    float* f32;
    __hip_bfloat16* hi;
    __hip_bfloat16* lo;  // hi + 8
    int prec;
    int nc;
    size_t cs;
};
__device__ inline size_t il(size_t x) { return 2 * (x & ~(size_t)7) + (x & 7); }

// c = exp(H_i) e_1 for every column: one wave per column, H in LDS.
// Scaling and squaring (the method of Eigen's MatrixExponential, which the
// reference's build_P calls): with ||H/2^s||_1 <= 1/2 each factor
// exp(H/2^s) is a Taylor polynomial of degree <= 24. When 2^s <= m the
// factor is applied 2^s times to the vector; otherwise E = exp(H/2^s) is
// formed column by column and squared s times, and c is its first column
// (cost O(s m^3) instead of O(2^s m^2) for large norms).
__device__ inline void kr_taylor_apply(const double* Hs, double* v, double* t, int m, int l) {
    double acc = l < m ? v[l] : 0.;
    double term = acc;
    for (int kk = 1; kk <= 24; kk++) {
        if (l < m) t[l] = term;
        __syncthreads();
        double nt = 0;
        if (l < m)
            for (int c = 0; c < m; c++) nt += Hs[l + c * m] * t[c];
        nt /= kk;
        __syncthreads();
        term = nt;
        acc += nt;
        double mag = fabs(nt);
        for (int d = 32; d > 0; d >>= 1) mag = fmax(mag, __shfl_xor(mag, d, 64));
        if (mag < 1e-18) break;
    }
    if (l < m) v[l] = acc;
    __syncthreads();
}

__global__ __launch_bounds__(64) void kr_expm_e1(const double* __restrict__ H, double* __restrict__ C,
                                                 int m) {
    extern __shared__ double sh[];
    double* Hs = sh;            // m*m
    double* v = sh + m * m;     // m
    double* t = v + m;          // m
    double* E = t + m;          // m*m (squaring path)
    double* W = E + m * m;      // m*m (squaring path)
    const int i = blockIdx.x, l = threadIdx.x;
    const double* Hi = H + (size_t)i * m * m;
    for (int q = l; q < m * m; q += 64) Hs[q] = Hi[q];
    __syncthreads();
    // 1-norm
    double cs = 0;
    if (l < m)
        for (int r = 0; r < m; r++) cs += fabs(Hs[r + l * m]);
    for (int d = 32; d > 0; d >>= 1) cs = fmax(cs, __shfl_xor(cs, d, 64));
    // a non-finite Hessenberg matrix yields NaN
    if (!(cs < 1.0e300)) {
        if (l < m) C[(size_t)i * m + l] = __builtin_nan("");
        return;
    }
    int s = 0;
    if (cs > 0.5) s = (int)ceil(log2(cs / 0.5));
    const double sc = ldexp(1.0, -s);
    for (int q = l; q < m * m; q += 64) Hs[q] *= sc;
    __syncthreads();
    if (s < 31 && (1L << s) <= m) {
        if (l < m) v[l] = l == 0 ? 1.0 : 0.0;
        __syncthreads();
        for (long rp = 0; rp < (1L << s); rp++) kr_taylor_apply(Hs, v, t, m, l);
        if (l < m) C[(size_t)i * m + l] = v[l];
        return;
    }
    for (int j = 0; j < m; j++) {
        if (l < m) v[l] = l == j ? 1.0 : 0.0;
        __syncthreads();
        kr_taylor_apply(Hs, v, t, m, l);
        if (l < m) E[l + j * m] = v[l];
    }
    __syncthreads();
    for (int q = 0; q < s; q++) {
        // thread l forms column l of E E
        if (l < m)
            for (int r = 0; r < m; r++) {
                double a = 0.;
                for (int c = 0; c < m; c++) a += E[r + c * m] * E[c + l * m];
                W[r + l * m] = a;
            }
        __syncthreads();
        for (int u = l; u < m * m; u += 64) E[u] = W[u];
        __syncthreads();
    }
    if (l < m) C[(size_t)i * m + l] = E[l];
}

// Wide variants: one 1024-thread workgroup per column, element r of the
// column held by thread r mod 1024 (E = ceil(N/1024) per thread), the basis
// type fixed at compile time, V_{j+1} prefetched while V_j's projection is
// reduced, one barrier per projection. Same operation order per column as
// Arnoldi_algorithm: each projection uses the updated vector.
template <int PREC>
__device__ inline double bload(const Basis& V, size_t i) {
    if constexpr (PREC == 0) return V.f64[i];
    else if constexpr (PREC == 1) return (double)V.f32[i];
    else return (double)__bfloat162float(V.hi[il(i)]) + (double)__bfloat162float(V.lo[il(i)]);
}
template <int PREC>
__device__ inline void bstore(const Basis& V, size_t i, double v) {
    if constexpr (PREC == 0) V.f64[i] = v;
    else if constexpr (PREC == 1) V.f32[i] = (float)v;
    else {
        float f = (float)v;
        __hip_bfloat16 h = __float2bfloat16(f);
        V.hi[il(i)] = h;
        V.lo[il(i)] = __float2bfloat16(f - __bfloat162float(h));
    }
}
// sum over the 16 waves of a 1024-thread workgroup; buf holds 16 doubles
// and must not be rewritten before every thread has read it
__device__ inline double wide_sum(double v, double* buf) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) s += buf[q];
    return s;
}

template <int E, int PREC>
__global__ __launch_bounds__(1024) void kr_mgs_wide(const float* __restrict__ Z32, const double* __restrict__ Z64,
                                                    Basis V, double* __restrict__ H, int* __restrict__ closed,
                                                    int N, size_t ldz, int m, int k, int textbook) {
    __shared__ double red[3][16];
    __shared__ double hz;
    const int i = blockIdx.x;
    if (closed[i]) return;
    const int tid = threadIdx.x;
    double z[E], v[E];
    const size_t zc = (size_t)i * ldz;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int r = tid + 1024 * e;
        z[e] = r < N ? (PREC == 0 ? Z64[zc + r] : (double)Z32[zc + r]) : 0.;
    }
    double* Hi = H + (size_t)i * m * m;  // column-major m x m
    const int jmax = textbook ? k + 1 : k;
    // j = 0: V_0(:,i) = e_i, so the projection is element i of z
#pragma unroll
    for (int e = 0; e < E; e++)
        if (tid + 1024 * e == i) hz = z[e];
    if (jmax > 1) {
        const size_t vc = ((size_t)1 * V.nc + i) * V.cs;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int r = tid + 1024 * e;
            v[e] = r < N ? bload<PREC>(V, vc + r) : 0.;
        }
    }
    __syncthreads();
    if (jmax > 0) {
        const double h = hz;
        if (tid == 0) Hi[(size_t)k * m] = h;
#pragma unroll
        for (int e = 0; e < E; e++)
            if (tid + 1024 * e == i) z[e] -= h;
    }
    for (int j = 1; j < jmax; j++) {
        double d = 0;
#pragma unroll
        for (int e = 0; e < E; e++) d += v[e] * z[e];
        double vn[E];
        if (j + 1 < jmax) {
            const size_t vc = ((size_t)(j + 1) * V.nc + i) * V.cs;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int r = tid + 1024 * e;
                vn[e] = r < N ? bload<PREC>(V, vc + r) : 0.;
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) vn[e] = 0.;
        }
        const double h = wide_sum(d, red[j & 1]);
        if (tid == 0) Hi[j + (size_t)k * m] = h;
#pragma unroll
        for (int e = 0; e < E; e++) {
            z[e] -= h * v[e];
            v[e] = vn[e];
        }
    }
    double nn = 0;
#pragma unroll
    for (int e = 0; e < E; e++) nn += z[e] * z[e];
    const double zn = sqrt(wide_sum(nn, red[2]));
    if (almost_equal(zn, 0., 1.0e-12, 1.0e-9)) {
        if (tid == 0) closed[i] = 1;
        return;
    }
    if (k + 1 < m) {
        if (tid == 0) Hi[(k + 1) + (size_t)k * m] = zn;
        const size_t vc = ((size_t)(k + 1) * V.nc + i) * V.cs;
        const double inv = 1.0 / zn;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int r = tid + 1024 * e;
            if (r < N) bstore<PREC>(V, vc + r, z[e] * inv);
        }
    }
}

template <int E, int PREC>
__global__ __launch_bounds__(1024) void kr_combine_wide(Basis V, const double* __restrict__ C,
                                                        double* __restrict__ P, double* __restrict__ colerr, int N,
                                                        int m, int project) {
    __shared__ double red[16];
    const int i = blockIdx.x, tid = threadIdx.x;
    const double* c = C + (size_t)i * m;
    double acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = tid + 1024 * e == i ? c[0] : 0.;
    for (int k = 1; k < m; k++) {
        const double ck = c[k];
        if (ck == 0.) continue;
        const size_t vc = ((size_t)k * V.nc + i) * V.cs;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int r = tid + 1024 * e;
            if (r < N) acc[e] += ck * bload<PREC>(V, vc + r);
        }
    }
    double s = 0;
#pragma unroll
    for (int e = 0; e < E; e++) s += acc[e];
    const double tot = wide_sum(s, red);
    if (tid == 0) colerr[i] = fabs(tot - 1.0);
    const double corr = project ? (1.0 - tot) / N : 0.;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int r = tid + 1024 * e;
        if (r < N) P[(size_t)i * N + r] = acc[e] + corr;
    }
}

// Vector variants (N a multiple of 8): NT threads per column, thread t owns
// the contiguous elements t*E .. t*E+E-1 (E a multiple of 8) and moves them
// with 16-byte loads and stores; chunks of 8 past N are skipped. One
// reduction per projection over NT/64 waves (the order of Arnoldi_algorithm:
// each projection uses the updated vector). About 2.5x the bandwidth of the
// wide variants at N = 8192 (fewer load instructions, 8 waves per column).
template <int PREC>
__device__ inline void vload8(const Basis& V, size_t i, double* o) {
    if constexpr (PREC == 0) {
        const double2* p = reinterpret_cast<const double2*>(V.f64 + i);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const double2 v = p[c];
            o[2 * c] = v.x;
            o[2 * c + 1] = v.y;
        }
    } else if constexpr (PREC == 1) {
        const float4* p = reinterpret_cast<const float4*>(V.f32 + i);
        const float4 a = p[0], b = p[1];
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
        o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
    } else {
        // i is a multiple of 8: 16 bytes of hi then 16 of lo at 2 i
        const uint4 h = *reinterpret_cast<const uint4*>(V.hi + 2 * i);
        const uint4 l = *reinterpret_cast<const uint4*>(V.lo + 2 * i);
        const uint32_t hh[4] = {h.x, h.y, h.z, h.w}, ll[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
        for (int w = 0; w < 4; w++) {
            // hi + lo of a bf16 pair is exact in FP32 (16 significant bits)
            o[2 * w] = (double)(__uint_as_float(hh[w] << 16) + __uint_as_float(ll[w] << 16));
            o[2 * w + 1] = (double)(__uint_as_float(hh[w] & 0xffff0000u) + __uint_as_float(ll[w] & 0xffff0000u));
        }
    }
}
template <int PREC>
__device__ inline void vstore8(const Basis& V, size_t i, const double* v) {
    if constexpr (PREC == 0) {
        double2* p = reinterpret_cast<double2*>(V.f64 + i);
#pragma unroll
        for (int c = 0; c < 4; c++) p[c] = make_double2(v[2 * c], v[2 * c + 1]);
    } else if constexpr (PREC == 1) {
        float4* p = reinterpret_cast<float4*>(V.f32 + i);
        p[0] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        p[1] = make_float4((float)v[4], (float)v[5], (float)v[6], (float)v[7]);
    } else {
        uint32_t hh[4], ll[4];
#pragma unroll
        for (int w = 0; w < 4; w++) {
            uint32_t hb[2], lb[2];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const float f = (float)v[2 * w + q];
                const __hip_bfloat16 h = __float2bfloat16(f);
                const __hip_bfloat16 l = __float2bfloat16(f - __bfloat162float(h));
                hb[q] = (uint32_t)__bfloat16_as_ushort(h);
                lb[q] = (uint32_t)__bfloat16_as_ushort(l);
            }
            hh[w] = hb[0] | (hb[1] << 16);
            ll[w] = lb[0] | (lb[1] << 16);
        }
        *reinterpret_cast<uint4*>(V.hi + 2 * i) = make_uint4(hh[0], hh[1], hh[2], hh[3]);
        *reinterpret_cast<uint4*>(V.lo + 2 * i) = make_uint4(ll[0], ll[1], ll[2], ll[3]);
    }
}
// sum over the NT/64 waves of the workgroup (buf: NT/64 doubles, not
// rewritten before every thread has read it)
template <int NT>
__device__ inline double vec_sum(double v, double* buf) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; q++) s += buf[q];
    return s;
}

template <int NT, int E, int PREC>
__global__ __launch_bounds__(NT) void kr_mgs_vec(const float* __restrict__ Z32, const double* __restrict__ Z64,
                                                 Basis V, double* __restrict__ H, int* __restrict__ closed, int N,
                                                 size_t ldz, int m, int k, int textbook) {
    __shared__ double red[3][NT / 64];
    const int i = blockIdx.x;
    if (closed[i]) return;
    const int r0 = threadIdx.x * E;
    double z[E];
#pragma unroll
    for (int c = 0; c < E / 8; c++) {
        const int r = r0 + 8 * c;
        if (r < N) {
            if (PREC == 0) {
                const double2* p = reinterpret_cast<const double2*>(Z64 + (size_t)i * ldz + r);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double2 v = p[q];
                    z[8 * c + 2 * q] = v.x;
                    z[8 * c + 2 * q + 1] = v.y;
                }
            } else {
                const float4* p = reinterpret_cast<const float4*>(Z32 + (size_t)i * ldz + r);
                const float4 a = p[0], b = p[1];
                z[8 * c] = a.x; z[8 * c + 1] = a.y; z[8 * c + 2] = a.z; z[8 * c + 3] = a.w;
                z[8 * c + 4] = b.x; z[8 * c + 5] = b.y; z[8 * c + 6] = b.z; z[8 * c + 7] = b.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) z[8 * c + q] = 0.;
        }
    }
    double* Hi = H + (size_t)i * m * m;  // column-major m x m
    const int jmax = textbook ? k + 1 : k;
    // j = 0: V_0(:,i) = e_i, so the projection is element i of z, which then
    // becomes 0; only the thread holding it takes part
    if (jmax > 0 && i >= r0 && i < r0 + E) {
#pragma unroll
        for (int e = 0; e < E; e++)
            if (r0 + e == i) {
                Hi[(size_t)k * m] = z[e];
                z[e] = 0.;
            }
    }
    for (int j = 1; j < jmax; j++) {
        double v[E];
        const size_t vc = ((size_t)j * V.nc + i) * V.cs;
#pragma unroll
        for (int c = 0; c < E / 8; c++) {
            if (r0 + 8 * c < N) vload8<PREC>(V, vc + r0 + 8 * c, v + 8 * c);
            else {
#pragma unroll
                for (int q = 0; q < 8; q++) v[8 * c + q] = 0.;
            }
        }
        double d = 0;
#pragma unroll
        for (int e = 0; e < E; e++) d += v[e] * z[e];
        const double h = vec_sum<NT>(d, red[j & 1]);
        if (threadIdx.x == 0) Hi[j + (size_t)k * m] = h;
#pragma unroll
        for (int e = 0; e < E; e++) z[e] -= h * v[e];
    }
    double nn = 0;
#pragma unroll
    for (int e = 0; e < E; e++) nn += z[e] * z[e];
    const double zn = sqrt(vec_sum<NT>(nn, red[2]));
    if (almost_equal(zn, 0., 1.0e-12, 1.0e-9)) {
        if (threadIdx.x == 0) closed[i] = 1;
        return;
    }
    if (k + 1 < m) {
        if (threadIdx.x == 0) Hi[(k + 1) + (size_t)k * m] = zn;
        const size_t vc = ((size_t)(k + 1) * V.nc + i) * V.cs;
        const double inv = 1.0 / zn;
#pragma unroll
        for (int c = 0; c < E / 8; c++) {
            if (r0 + 8 * c >= N) continue;
            double o[8];
#pragma unroll
            for (int q = 0; q < 8; q++) o[q] = z[8 * c + q] * inv;
            vstore8<PREC>(V, vc + r0 + 8 * c, o);
        }
    }
}

template <int NT, int E, int PREC>
__global__ __launch_bounds__(NT) void kr_combine_vec(Basis V, const double* __restrict__ C, double* __restrict__ P,
                                                     double* __restrict__ colerr, int N, int m, int project) {
    __shared__ double red[NT / 64];
    const int i = blockIdx.x, r0 = threadIdx.x * E;
    const double* c = C + (size_t)i * m;
    double acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = r0 + e == i ? c[0] : 0.;
    for (int k = 1; k < m; k++) {
        const double ck = c[k];
        if (ck == 0.) continue;
        const size_t vc = ((size_t)k * V.nc + i) * V.cs;
#pragma unroll
        for (int q = 0; q < E / 8; q++) {
            if (r0 + 8 * q >= N) continue;
            double v[8];
            vload8<PREC>(V, vc + r0 + 8 * q, v);
#pragma unroll
            for (int e = 0; e < 8; e++) acc[8 * q + e] += ck * v[e];
        }
    }
    double s = 0;
#pragma unroll
    for (int e = 0; e < E; e++) s += acc[e];
    const double tot = vec_sum<NT>(s, red);
    if (threadIdx.x == 0) colerr[i] = fabs(tot - 1.0);
    const double corr = project ? (1.0 - tot) / N : 0.;
#pragma unroll
    for (int q = 0; q < E / 8; q++) {
        if (r0 + 8 * q >= N) continue;
        double2* p = reinterpret_cast<double2*>(P + (size_t)i * N + r0 + 8 * q);
#pragma unroll
        for (int e = 0; e < 4; e++) p[e] = make_double2(acc[8 * q + 2 * e] + corr, acc[8 * q + 2 * e + 1] + corr);
    }
}

// launch shape of the vector variants: (threads, elements per thread)
static void vec_shape(int N, int& nt, int& e) {
    if (N <= 2048) { nt = 256; e = 8; }
    else if (N <= 4096) { nt = 512; e = 8; }
    else if (N <= 8192) { nt = 512; e = 16; }
    else { nt = 1024; e = 16; }
}
template <int NT, int E>
static void launch_vec(hipStream_t s, int N, size_t ldz, const float* z32, const double* z64, Basis V, double* H,
                       int* closed, int m, int k, int tb) {
    if (V.prec == 0)
        hipLaunchKernelGGL((kr_mgs_vec<NT, E, 0>), dim3(N), dim3(NT), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
    else if (V.prec == 1)
        hipLaunchKernelGGL((kr_mgs_vec<NT, E, 1>), dim3(N), dim3(NT), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
    else
        hipLaunchKernelGGL((kr_mgs_vec<NT, E, 2>), dim3(N), dim3(NT), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
}
template <int NT, int E>
static void launch_combine_vec(hipStream_t s, int N, Basis V, const double* C, double* P, double* err, int m) {
    if (V.prec == 0)
        hipLaunchKernelGGL((kr_combine_vec<NT, E, 0>), dim3(N), dim3(NT), 0, s, V, C, P, err, N, m, 1);
    else if (V.prec == 1)
        hipLaunchKernelGGL((kr_combine_vec<NT, E, 1>), dim3(N), dim3(NT), 0, s, V, C, P, err, N, m, 1);
    else
        hipLaunchKernelGGL((kr_combine_vec<NT, E, 2>), dim3(N), dim3(NT), 0, s, V, C, P, err, N, m, 1);
}
static void mgs_step(hipStream_t s, int N, size_t ldz, const float* z32, const double* z64, Basis V, double* H,
                     int* closed, int m, int k, int tb) {
    int nt, e;
    vec_shape(N, nt, e);
    if (nt == 256) launch_vec<256, 8>(s, N, ldz, z32, z64, V, H, closed, m, k, tb);
    else if (nt == 512 && e == 8) launch_vec<512, 8>(s, N, ldz, z32, z64, V, H, closed, m, k, tb);
    else if (nt == 512) launch_vec<512, 16>(s, N, ldz, z32, z64, V, H, closed, m, k, tb);
    else launch_vec<1024, 16>(s, N, ldz, z32, z64, V, H, closed, m, k, tb);
}
static void combine_step(hipStream_t s, int N, Basis V, const double* C, double* P, double* err, int m) {
    int nt, e;
    vec_shape(N, nt, e);
    if (nt == 256) launch_combine_vec<256, 8>(s, N, V, C, P, err, m);
    else if (nt == 512 && e == 8) launch_combine_vec<512, 8>(s, N, V, C, P, err, m);
    else if (nt == 512) launch_combine_vec<512, 16>(s, N, V, C, P, err, m);
    else launch_combine_vec<1024, 16>(s, N, V, C, P, err, m);
}

template <int E>
static void launch_mgs_wide(hipStream_t s, int N, size_t ldz, const float* z32, const double* z64, Basis V,
                            double* H, int* closed, int m, int k, int tb) {
    if (V.prec == 0)
        hipLaunchKernelGGL((kr_mgs_wide<E, 0>), dim3(N), dim3(1024), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
    else if (V.prec == 1)
        hipLaunchKernelGGL((kr_mgs_wide<E, 1>), dim3(N), dim3(1024), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
    else
        hipLaunchKernelGGL((kr_mgs_wide<E, 2>), dim3(N), dim3(1024), 0, s, z32, z64, V, H, closed, N, ldz, m, k, tb);
}
template <int E>
static void launch_combine_wide(hipStream_t s, int N, Basis V, const double* C, double* P, double* err, int m) {
    if (V.prec == 0)
        hipLaunchKernelGGL((kr_combine_wide<E, 0>), dim3(N), dim3(1024), 0, s, V, C, P, err, N, m, 1);
    else if (V.prec == 1)
        hipLaunchKernelGGL((kr_combine_wide<E, 1>), dim3(N), dim3(1024), 0, s, V, C, P, err, N, m, 1);
    else
        hipLaunchKernelGGL((kr_combine_wide<E, 2>), dim3(N), dim3(1024), 0, s, V, C, P, err, N, m, 1);
}
static int wide_elems(int N) { return N <= 1024 ? 1 : (N <= 2048 ? 2 : (N <= 4096 ? 4 : (N <= 8192 ? 8 : 16))); }


}  // namespace bte

using namespace bte;

namespace bte {
// rocBLAS handle of a context, created on first use and bound to its stream
void* blas_handle(Ctx* c) {
    if (!c->blas) {
        rocblas_handle h0;
        RBCHECK(rocblas_create_handle(&h0));
        c->blas = h0;
        c->blas_free = [](void* h) { rocblas_destroy_handle((rocblas_handle)h); };
    }
    RBCHECK(rocblas_set_stream((rocblas_handle)c->blas, c->stream));
    return c->blas;
}
}  // namespace bte

extern "C" int bte_krylov_build(bte_ctx* ctx, int N, const double* B, double dt, int m, int prec,
                                int mode, double* P, bte_krylov_report* rep) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !B || !P || N < 2 || m < 2 || m > 64 || prec < 0 || prec > 2 || (mode != 0 && mode != 1))
        throw Error(BTE_EINVAL, "bad arguments to bte_krylov_build");
    if (N > 16384) throw Error(BTE_EINVAL, "N above 16384 is not supported");
    const int WE = wide_elems(N);
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t NN = (size_t)N * N;
    rocblas_handle hb = (rocblas_handle)blas_handle(c);
    DevBuf<double> dB, dZ64, dH, dC, dP, dErr;
    DevBuf<float> dA32, dZ32, dV32;
    DevBuf<double> dV64;
    DevBuf<__hip_bfloat16> dAil, dVil;
    DevBuf<int> dClosed;
    dB.upload(B, NN, s);
    Basis V{nullptr, nullptr, nullptr, nullptr, prec, N, (size_t)N};
    // BF16: N padded to the product kernel's 256 with zero rows and columns
    const int Np = prec == 2 ? (N + gk::TM - 1) / gk::TM * gk::TM : N;
    const size_t ldz = (size_t)Np;
    const size_t lda_il = 2 * (size_t)Np + 64;  // bf16 per row of A (the pad spreads rows over L2 channels)
    const int cb = 4 * c->cu_count;
    // A = dt*B in the compute format; Z_0 = dt*B*V_0 = dt*B
    if (prec == 0) {
        dZ64.alloc(NN);
        dV64.alloc(NN * m);
        V.f64 = dV64.p;
        hipLaunchKernelGGL(kr_scale64, dim3(cb), dim3(256), 0, s, dB.p, dZ64.p, NN, dt);
    } else if (prec == 1) {
        dA32.alloc(NN);
        dZ32.alloc(NN);
        dV32.alloc(NN * m);
        V.f32 = dV32.p;
        hipLaunchKernelGGL(kr_to_f32, dim3(cb), dim3(256), 0, s, dB.p, dA32.p, NN, dt);
        HIPCHECK(hipMemcpyAsync(dZ32.p, dA32.p, NN * sizeof(float), hipMemcpyDeviceToDevice, s));
    } else {
        const size_t NNp = (size_t)Np * Np;
        dAil.alloc((size_t)Np * lda_il);
        dZ32.alloc(NNp);
        V.nc = Np;
        V.cs = (size_t)Np + 32;
        dVil.alloc((size_t)m * Np * 2 * V.cs);
        V.hi = dVil.p;
        V.lo = dVil.p + 8;
        hipLaunchKernelGGL(kr_to_bf16_il, dim3(Np / 32, Np / 32), dim3(256), 0, s, dB.p, dAil.p, N, lda_il, dt);
        hipLaunchKernelGGL(kr_to_f32_ld, dim3(cb), dim3(256), 0, s, dB.p, dZ32.p, N, ldz, dt);
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&kr_bf16x3_gemm),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, gk::LDS));
    }
    HIPCHECK(hipGetLastError());
    dH.alloc((size_t)N * m * m);
    dC.alloc((size_t)N * m);
    dP.alloc(NN);
    dErr.alloc(N);
    dClosed.alloc(N);
    HIPCHECK(hipMemsetAsync(dH.p, 0, (size_t)N * m * m * sizeof(double), s));
    HIPCHECK(hipMemsetAsync(dClosed.p, 0, N * sizeof(int), s));
    if (prec != 0) dB.release();
    // columns of a basis that closes early stay zero
    if (prec == 0) HIPCHECK(hipMemsetAsync(dV64.p, 0, NN * m * sizeof(double), s));
    if (prec == 1) HIPCHECK(hipMemsetAsync(dV32.p, 0, NN * m * sizeof(float), s));
    // BF16: also the padding rows and columns, which the product reads
    if (prec == 2) HIPCHECK(hipMemsetAsync(dVil.p, 0, (size_t)m * Np * 2 * V.cs * sizeof(__hip_bfloat16), s));
    // events bracket the build and every GEMM; they are read after the final
    // synchronisation, so the host never waits inside the loop
    struct EvGuard {
        std::vector<hipEvent_t> ev;
        ~EvGuard() {
            for (auto e : ev) (void)hipEventDestroy(e);
        }
    } eg;
    eg.ev.resize(2 + 2 * (size_t)m);
    for (auto& e : eg.ev) HIPCHECK(hipEventCreate(&e));
    const hipEvent_t e0 = eg.ev[0], e1 = eg.ev[1];
    HIPCHECK(hipEventRecord(e0, s));
    int gemms = 0;
    const bool vec = N % 8 == 0;  // vector Gram-Schmidt and combine kernels
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    for (int k = 0; k < m; k++) {
        if (k > 0) {
            // Z = (dt*B) * V_k
            HIPCHECK(hipEventRecord(eg.ev[2 + 2 * gemms], s));
            const size_t vo = (size_t)k * NN;
            if (prec == 0) {
                RBCHECK(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, N, N, N, &dt, dB.p, N,
                                      dV64.p + vo, N, &zero, dZ64.p, N));
            } else if (prec == 1) {
                RBCHECK(rocblas_sgemm(hb, rocblas_operation_none, rocblas_operation_none, N, N, N, &onef, dA32.p,
                                      N, dV32.p + vo, N, &zerof, dZ32.p, N));
            } else {
                // A_hi V_hi + A_hi V_lo + A_lo V_hi on MFMA; V_k's columns
                // start at logical (k Np) cs, stored at twice that
                hipLaunchKernelGGL(kr_bf16x3_gemm, dim3((Np / gk::TM) * (Np / gk::TN)), dim3(gk::NT), gk::LDS, s,
                                   dAil.p, lda_il, dVil.p + 2 * (size_t)k * Np * V.cs, 2 * V.cs, dZ32.p, Np);
            }
            HIPCHECK(hipEventRecord(eg.ev[3 + 2 * gemms], s));
            gemms++;
        }
        const float* z32 = prec == 0 ? nullptr : dZ32.p;
        const double* z64 = prec == 0 ? dZ64.p : nullptr;
        if (vec) {
            mgs_step(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode);
        } else switch (WE) {
            case 1: launch_mgs_wide<1>(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode); break;
            case 2: launch_mgs_wide<2>(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode); break;
            case 4: launch_mgs_wide<4>(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode); break;
            case 8: launch_mgs_wide<8>(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode); break;
            default: launch_mgs_wide<16>(s, N, ldz, z32, z64, V, dH.p, dClosed.p, m, k, mode); break;
        }
        HIPCHECK(hipGetLastError());
    }
    {
        const size_t lds = (size_t)(3 * m * m + 2 * m) * sizeof(double);
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&kr_expm_e1),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kr_expm_e1, dim3(N), dim3(64), lds, s, dH.p, dC.p, m);
    }
    HIPCHECK(hipGetLastError());
    if (vec) {
        combine_step(s, N, V, dC.p, dP.p, dErr.p, m);
    } else switch (WE) {
        case 1: launch_combine_wide<1>(s, N, V, dC.p, dP.p, dErr.p, m); break;
        case 2: launch_combine_wide<2>(s, N, V, dC.p, dP.p, dErr.p, m); break;
        case 4: launch_combine_wide<4>(s, N, V, dC.p, dP.p, dErr.p, m); break;
        case 8: launch_combine_wide<8>(s, N, V, dC.p, dP.p, dErr.p, m); break;
        default: launch_combine_wide<16>(s, N, V, dC.p, dP.p, dErr.p, m); break;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(e1, s));
    HIPCHECK(hipMemcpyAsync(P, dP.p, NN * sizeof(double), hipMemcpyDeviceToHost, s));
    std::vector<double> err(N);
    std::vector<int> closed(N);
    HIPCHECK(hipMemcpyAsync(err.data(), dErr.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(closed.data(), dClosed.p, N * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    double gemm_ms = 0;
    for (int q = 0; q < gemms; q++) {
        float gm = 0;
        HIPCHECK(hipEventElapsedTime(&gm, eg.ev[2 + 2 * q], eg.ev[3 + 2 * q]));
        gemm_ms += gm;
    }
    if (rep) {
        rep->build_ms = ms;
        rep->gemm_ms = gemm_ms;
        // algorithmic rate: 2 N^3 per product dt*B*V_k, whatever the number
        // of BF16 partial products that form it
        double flops = 2.0 * (double)N * N * N * gemms;
        rep->gemm_tflops = gemm_ms > 0 ? flops / (gemm_ms * 1e-3) / 1e12 : 0;
        double mx = 0;
        for (double e : err) mx = std::max(mx, e);
        rep->max_col_sum_err = mx;
        int nb = 0;
        for (int cl : closed) nb += cl;
        rep->breakdowns = nb;
    }
    BTE_GUARD_END
}

// ---------------------------------------------------------------------------
// FP64 reference exponential exp(t A): Pade(13) scaling and squaring
// (Higham, SIAM J. Matrix Anal. Appl. 26 (2005) 1179), the degree-13 branch
// for every norm, with the matrix products as rocBLAS FP64 GEMMs and the
// rational solve as a rocSOLVER LU factorisation. Column-major N x N.
namespace bte {
// out = (acc ? out : 0) + c6 A6 + c4 A4 + c2 A2 + cI I
__global__ __launch_bounds__(256) void ex_comb_kernel(double* out, const double* A6, const double* A4,
                                                      const double* A2, double c6, double c4, double c2, double cI,
                                                      int n, int acc) {
    const size_t nn = (size_t)n * n;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (size_t)gridDim.x * blockDim.x) {
        double v = c6 * A6[q] + c4 * A4[q] + c2 * A2[q];
        if (q % (size_t)(n + 1) == 0) v += cI;
        out[q] = acc ? out[q] + v : v;
    }
}
// a = s a;  d = u - v, e = u + v (for the solve)
__global__ __launch_bounds__(256) void ex_scale_kernel(double* a, double s, size_t nn) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (size_t)gridDim.x * blockDim.x)
        a[q] *= s;
}
__global__ __launch_bounds__(256) void ex_pm_kernel(double* vmu, double* vpu, const double* u, size_t nn) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (size_t)gridDim.x * blockDim.x) {
        const double v = vpu[q], uu = u[q];
        vmu[q] = v - uu;
        vpu[q] = v + uu;
    }
}
}  // namespace bte

extern "C" int bte_expm(bte_ctx* ctx, int N, const double* A, double t, double* E, double* ms) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !A || !E || N < 1) throw Error(BTE_EINVAL, "bad arguments to bte_expm");
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const size_t nn = (size_t)N * N;
    static const double b[14] = {64764752532480000., 32382376266240000., 7771770303897600., 1187353796428800.,
                                 129060195264000., 10559470521600., 670442572800., 33522128640., 1323241920.,
                                 40840800., 960960., 16380., 182., 1.};
    const double theta13 = 5.371920351148152;
    double norm1 = 0;
    for (int j = 0; j < N; j++) {
        double sum = 0;
        for (int i = 0; i < N; i++) sum += std::fabs(t * A[i + (size_t)j * N]);
        norm1 = std::max(norm1, sum);
    }
    const int sq = norm1 > theta13 ? (int)std::ceil(std::log2(norm1 / theta13)) : 0;
    DevBuf<double> X, A2, A4, A6, T, U, V;
    DevBuf<rocblas_int> piv, info;
    X.upload(A, nn, s);
    A2.alloc(nn); A4.alloc(nn); A6.alloc(nn); T.alloc(nn); U.alloc(nn); V.alloc(nn);
    piv.alloc(N); info.alloc(1);
    const int gb = 4 * c->cu_count;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s));
    hipLaunchKernelGGL(ex_scale_kernel, dim3(gb), dim3(256), 0, s, X.p, std::ldexp(t, -sq), nn);
    rocblas_handle hb = (rocblas_handle)blas_handle(c);
    const double one = 1.0, zero = 0.0;
    auto mm = [&](const double* a, const double* bb, double* out) {
        RBCHECK(rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, N, N, N, &one, a, N, bb, N, &zero,
                              out, N));
    };
    mm(X.p, X.p, A2.p);
    mm(A2.p, A2.p, A4.p);
    mm(A4.p, A2.p, A6.p);
    // U = X (A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I)
    hipLaunchKernelGGL(ex_comb_kernel, dim3(gb), dim3(256), 0, s, T.p, A6.p, A4.p, A2.p, b[13], b[11], b[9], 0., N, 0);
    mm(A6.p, T.p, V.p);
    hipLaunchKernelGGL(ex_comb_kernel, dim3(gb), dim3(256), 0, s, V.p, A6.p, A4.p, A2.p, b[7], b[5], b[3], b[1], N, 1);
    mm(X.p, V.p, U.p);
    // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
    hipLaunchKernelGGL(ex_comb_kernel, dim3(gb), dim3(256), 0, s, T.p, A6.p, A4.p, A2.p, b[12], b[10], b[8], 0., N, 0);
    mm(A6.p, T.p, V.p);
    hipLaunchKernelGGL(ex_comb_kernel, dim3(gb), dim3(256), 0, s, V.p, A6.p, A4.p, A2.p, b[6], b[4], b[2], b[0], N, 1);
    // (V - U) R = V + U
    hipLaunchKernelGGL(ex_pm_kernel, dim3(gb), dim3(256), 0, s, T.p, V.p, U.p, nn);
    RBCHECK(rocsolver_dgetrf(hb, N, N, T.p, N, piv.p, info.p));
    RBCHECK(rocsolver_dgetrs(hb, rocblas_operation_none, N, N, T.p, N, piv.p, V.p, N));
    double* cur = V.p;
    double* nxt = U.p;
    for (int k = 0; k < sq; k++) {
        mm(cur, cur, nxt);
        std::swap(cur, nxt);
    }
    HIPCHECK(hipEventRecord(e1, s));
    rocblas_int inf = 0;
    HIPCHECK(hipMemcpyAsync(E, cur, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float el = 0;
    HIPCHECK(hipEventElapsedTime(&el, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = el;
    if (inf != 0) throw Error(BTE_EINVAL, "singular Pade denominator in bte_expm");
    BTE_GUARD_END
}
