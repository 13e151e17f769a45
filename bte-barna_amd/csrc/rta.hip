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
#include <tuple>
#include <limits>
#include <mutex>
#include <rocblas/rocblas.h>

namespace bte {

struct RtaGen {
    double p1x, p1y, ndx, ndy;
    int32_t idi, sign, pad0, pad1;
};

// One run of a batch: its particle-id range, seed, tables and options. All
// box, generator and table indices are absolute in the batch's arrays.
struct RtaRunDev {
    uint64_t first, count, seed;
    int32_t gen0, ngen, gsel0, scat0;
    int32_t ballistic, gmask, pad0, pad1;  // gmask: sides whose specularity gets a gradient
    double wscale;                         // scattering-rate factor of the run
    float spec[4];                         // specular probability per side
};

// Tally mode 3: a contiguous range of one run's particle ids, taken whole by
// one workgroup, and the run's rows in the workgroup's FP64 LDS block.
struct RtaItem {
    uint64_t lo, hi;     // particle ids [lo, hi)
    int32_t lrow0;       // global LDS row of the run's first regular box
    int32_t lrows;       // the run's regular boxes
    int32_t lbins;       // leading time bins of each row held in LDS
    int32_t run;
};

struct RtaParams {
    const double4* modes;  // (vx, vy, w, omega)
    const double2* ivel;   // (1/vx, 1/vy), 0 where the component is 0
    const int32_t* mirror; // [2][n]: partner with vx -> -vx (sides 0, 3), vy -> -vy (sides 1, 2)
    const uint2* scat;     // BE_derivative alias per run (n each)
    const DevBox* boxes;
    const DevContact* cts;
    const double4* cseg;    // contact segments (periodic translation)
    const double2* ptrans;  // periodic translation per box
    // polygon geometries (POLY kernels): sides, side range and centre per box
    const DevSide* sides;
    const int2* srange;
    const double2* center;
    const RtaGen* gens;
    const uint2* gen_sel;  // alias over each run's generators
    const uint2* gen_tab;  // per generator, n entries
    const int4* lam_info;  // per mode: lo, len, base, 0
    const uint2* lam_tab;
    const int32_t* perm;   // omega-sorted position -> mode
    const double* timegrid;
    const uint64_t* box_cell;  // first tally cell of each box (box-major, nt cells)
    const RtaRunDev* runs;
    double* gz;
    double* jx;
    double* jy;
    double* dgz;           // [4][cells] score-weighted gz per side (gradient runs)
    const int32_t* box_loc; // spectral location of each box, -1 none
    double* spec;          // [nloc][nt][n] per-mode gz of the spectral locations (null: none)
    unsigned long long* counters;  // events, crossings, next id, stuck, bad mirror, bad translation
    uint64_t total;        // particles over all runs
    uint64_t cells;
    unsigned long long max_iters;
    double maxtime;
    int32_t n, nt, nruns, pad;
    // FP32 tallies in LDS (LT kernels): the LDS row (nt cells) of each box,
    // ~0u for boxes that never tally; the global cell of each LDS row; per
    // side the float offset of its score-weighted array, -1 for none. The
    // LDS block holds 3 arrays (gz, jx, jy) of lrows * nt floats, then the
    // gradient arrays.
    const uint32_t* box_lrow;
    const uint64_t* lrow_cell;
    int32_t lrows, pad1;
    int32_t lt_goff[4];
    // tally mode 2: the LDS holds FP64 gz, jx, jy of the first lbins time
    // bins of every LDS row ([3][lrows][lbins]); each workgroup writes its
    // block to part[blockIdx.x] at the end
    int32_t lbins, pad2;
    double* part;
    double idt;  // nt / maxtime: time -> bin estimate on the uniform grid
    // tally mode 3: the work items, taken in order through counters[6]
    const RtaItem* items;
    int32_t nitems, pad3;
};

// Per-XCD copies of the f64 tallies. Device-wide f64 atomics to a few
// thousand cells are the RTAMC2D kernel's main cost; spread over one copy
// per XCD they run about 25 % faster.
#define RTA_NCOPY 8
// the XCD (XCC) this wave runs on, from the hardware register HW_REG_XCC_ID

// get_bin_index (RTA_MC2D.cpp:140-148): lower_bound - 1 clamped to
// [0, nt-1]. The grid is uniform (t_i = i maxtime / nt), so the bin is
// estimated from target * nt / maxtime and corrected against the grid
// values themselves, which keeps lower_bound's result exactly (a target
// on a grid point t_j belongs to bin j - 1).
__device__ inline int time_bin(const double* grid, int nt, double idt, double target) {
    const double e = target * idt;
    int k = e < 0. ? 0 : (e > (double)nt ? nt : (int)e);
    while (k > 0 && grid[k] >= target) k--;
    while (k < nt && grid[k + 1] < target) k++;
    return k > nt - 1 ? nt - 1 : k;
}

// A lane's running sum of consecutive contributions to one tally cell
// (box, time bin); it is added to the tallies when the lane's next
// contribution falls in another cell and when the lane's histories end.
// Successive segments of a history often share a cell (several flights in
// one box within one time bin), and each cell update costs three adds.
struct TallyAcc {
    int box = -1, bin = 0;
    double g = 0., x = 0., y = 0.;
};

// Adds the accumulator into the tallies: under tally mode 2 the first
// lbins time bins of a box go to the workgroup's FP64 LDS block, later
// bins (and every bin under mode 0) to the XCD's copy by device atomics.
template <int TM>
__device__ inline void acc_flush(const RtaParams& P, double* L, TallyAcc& a) {
    if (a.box < 0) return;
    // mode 3: the block holds [3][lrows][lbins] of the current item's run,
    // after a header of ints (lrows, lbins, first global row)
    const int* hd = reinterpret_cast<const int*>(L) - 4;
    if (TM == 3 && a.bin < hd[1]) {
        const uint32_t lc = (uint32_t)hd[0] * (uint32_t)hd[1];
        const uint32_t i = (P.box_lrow[a.box] - (uint32_t)hd[2]) * (uint32_t)hd[1] + (uint32_t)a.bin;
        atomicAdd(L + i, a.g);
        atomicAdd(L + lc + i, a.x);
        atomicAdd(L + 2 * lc + i, a.y);
    } else if (TM == 2 && a.bin < P.lbins) {
        const uint32_t lc = (uint32_t)P.lrows * (uint32_t)P.lbins;
        const uint32_t i = P.box_lrow[a.box] * (uint32_t)P.lbins + (uint32_t)a.bin;
        atomicAdd(L + i, a.g);
        atomicAdd(L + lc + i, a.x);
        atomicAdd(L + 2 * lc + i, a.y);
    } else {
        const uint64_t k = P.box_cell[a.box] + (uint64_t)a.bin + (uint64_t)(xcd_id() % RTA_NCOPY) * 3 * P.cells;
        atomicAdd(P.gz + k, a.g);
        atomicAdd(P.jx + k, a.x);
        atomicAdd(P.jy + k, a.y);
    }
    a.box = -1;
}

// process_segment (RTA_MC2D.cpp:699-769); gradient runs also add the
// segment weighted by the particle's running score for each active side.
// TM (tally mode) 0: FP64 device atomics. 1: the workgroup's FP32 partial
// sums in LDS (lds), added into the FP64 tallies once at the end of the
// kernel (lt_flush). 2: the workgroup's FP64 partial sums of gz, jx and jy
// for the first lbins time bins in LDS, later bins by device atomics.
// 3: as 2 with the LDS block holding only the rows of the run of the
// workgroup's current work item. Modes 0 (without dT/dp), 2 and 3 merge
// consecutive contributions to one cell in the lane's accumulator.
template <bool GRAD, int TM>
__device__ inline void tally_segment(const RtaParams& P, void* lds, TallyAcc& acc, int box, double t0, double t1,
                                     int sign, double vx, double vy, int gmask, const float* score, int mode) {
    int b0 = time_bin(P.timegrid, P.nt, P.idt, t0), b1 = time_bin(P.timegrid, P.nt, P.idt, t1);
    double delta = t1 - t0;
    if constexpr (TM == 2 || TM == 3 || (TM == 0 && !GRAD)) {
        // spectral locations tally per mode; the Gamma-broadened spectra are
        // one matrix product after the run (the reference adds the broadened
        // weights of every segment, RTA_MC2D.cpp:718-766)
        double* sp = nullptr;
        if (TM == 0 && P.spec) {
            const int l = P.box_loc[box];
            if (l >= 0) sp = P.spec + (size_t)l * P.nt * P.n + mode;
        }
        for (int i = b0; i <= b1; i++) {
            double lo = fmax(P.timegrid[i] - t0, 0.);
            double hi = fmin(P.timegrid[i + 1] - t0, delta);
            double c = sign * fmax(hi - lo, 0.);
            if (c == 0.) continue;
            if (TM == 0 && sp) atomicAdd(sp + (size_t)i * P.n, c);
            if (box == acc.box && i == acc.bin) {
                acc.g += c;
                acc.x += vx * c;
                acc.y += vy * c;
            } else {
                acc_flush<TM>(P, (double*)lds, acc);
                acc.box = box;
                acc.bin = i;
                acc.g = c;
                acc.x = vx * c;
                acc.y = vy * c;
            }
        }
        return;
    }
    if constexpr (TM == 1) {
        float* lt = (float*)lds;
        const uint32_t lc = (uint32_t)P.lrows * (uint32_t)P.nt;
        float* L = lt + P.box_lrow[box] * (uint32_t)P.nt;
        for (int i = b0; i <= b1; i++) {
            double lo = fmax(P.timegrid[i] - t0, 0.);
            double hi = fmin(P.timegrid[i + 1] - t0, delta);
            double c = sign * fmax(hi - lo, 0.);
            if (c == 0.) continue;
            atomicAdd(L + i, (float)c);
            atomicAdd(L + lc + i, (float)(vx * c));
            atomicAdd(L + 2 * lc + i, (float)(vy * c));
            if (GRAD) {
#pragma unroll
                for (int s = 0; s < 4; s++)
                    if (P.lt_goff[s] >= 0 && ((gmask >> s) & 1) && score[s] != 0.f)
                        atomicAdd(L + P.lt_goff[s] + i, (float)(c * (double)score[s]));
            }
        }
        return;
    }
    // each XCD adds into its own copy of the tallies (RTA_NCOPY copies,
    // summed after the kernel): the copy is picked by the hardware XCD id,
    // so no two XCDs ever update one copy
    const uint64_t xc = xcd_id() % RTA_NCOPY;
    const uint64_t base = P.box_cell[box] + xc * 3 * P.cells;
    // spectral locations tally per mode; the Gamma-broadened spectra are one
    // matrix product after the run (the reference adds the broadened weights
    // of every segment, RTA_MC2D.cpp:718-766)
    double* sp = nullptr;
    if (P.spec) {
        const int l = P.box_loc[box];
        if (l >= 0) sp = P.spec + (size_t)l * P.nt * P.n + mode;
    }
    for (int i = b0; i <= b1; i++) {
        double lo = fmax(P.timegrid[i] - t0, 0.);
        double hi = fmin(P.timegrid[i + 1] - t0, delta);
        double c = sign * fmax(hi - lo, 0.);
        if (c == 0.) continue;
        const uint64_t k = base + i;
        atomicAdd(P.gz + k, c);
        atomicAdd(P.jx + k, vx * c);
        atomicAdd(P.jy + k, vy * c);
        if (sp) atomicAdd(sp + (size_t)i * P.n, c);
        if (GRAD) {
            for (int s = 0; s < 4; s++)
                if (((gmask >> s) & 1) && score[s] != 0.f)
                    atomicAdd(P.dgz + xc * 4 * P.cells + (size_t)s * P.cells + (k - xc * 3 * P.cells),
                              c * (double)score[s]);
        }
    }
}

// the run holding global particle id pid (first ids ascending)
__device__ inline int find_run(const RtaParams& P, uint64_t pid) {
    int lo = 0, hi = P.nruns - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (P.runs[mid].first <= pid) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The workgroup's LDS partial sums into the FP64 tallies (the XCD's copy).
template <bool GRAD>
__device__ inline void lt_flush(const RtaParams& P, const float* lt) {
    const uint32_t nt = (uint32_t)P.nt, lc = (uint32_t)P.lrows * nt;
    const uint64_t xc = xcd_id() % RTA_NCOPY;
    double* tz = P.gz + xc * 3 * P.cells;
    for (uint32_t j = threadIdx.x; j < lc; j += blockDim.x) {
        const uint32_t row = j / nt;
        const uint64_t g = P.lrow_cell[row] + (j - row * nt);
        for (int a = 0; a < 3; a++) {
            const float v = lt[(size_t)a * lc + j];
            if (v != 0.f) atomicAdd(tz + (uint64_t)a * P.cells + g, (double)v);
        }
        if (GRAD) {
            for (int s = 0; s < 4; s++) {
                if (P.lt_goff[s] < 0) continue;
                const float v = lt[(size_t)P.lt_goff[s] + j];
                if (v != 0.f) atomicAdd(P.dgz + xc * 4 * P.cells + (uint64_t)s * P.cells + g, (double)v);
            }
        }
    }
}

// The LT kernels give the whole LDS to the tallies (up to 160 KB), so their
// counters are summed through that block after the flush instead of
// block_add_u64's static array.
extern __shared__ __align__(16) float rta_lt[];
__device__ inline void lt_counters(unsigned long long* cnt, unsigned long long v0, unsigned long long v1,
                                   unsigned long long v3, unsigned long long v4, unsigned long long v5) {
    unsigned long long* part = (unsigned long long*)rta_lt;
    unsigned long long v[5] = {v0, v1, v3, v4, v5};
    for (int k = 0; k < 5; k++)
        for (int d = 32; d > 0; d >>= 1) v[k] += __shfl_down(v[k], d, 64);
    __syncthreads();  // the flush has read every partial sum
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; k++) part[(threadIdx.x >> 6) * 5 + k] = v[k];
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w * 5 + threadIdx.x];
        const int dst[5] = {0, 1, 3, 4, 5};
        if (t) atomicAdd(cnt + dst[threadIdx.x], t);
    }
}

// The state of one particle history between events.
struct RtaState {
    Pcg32 rng;
    double x, y, t, dtrem, wscale, tlast;
    unsigned long long budget;
    int mode, box, sign, run, still, alive, need_flight, scat0;
    int ballistic, pad0;
    float score[4];
};

struct RtaCounts {
    unsigned long long events = 0, crossings = 0, stuck = 0, badmirror = 0, badtrans = 0;
};

// Init_generator (RTA_MC2D.cpp:436-484): particle pid of its run at t = 0
template <bool GRAD, bool POLY>
__device__ inline void rta_start(const RtaParams& P, RtaState& S, uint64_t pid) {
    S.budget = 0;
    S.still = 0;
    S.tlast = -1.0;
    S.run = find_run(P, pid);
    const RtaRunDev& R = P.runs[S.run];
    S.rng.seed(R.seed, pid - R.first);
    S.wscale = R.wscale;
    S.scat0 = R.scat0;
    S.ballistic = R.ballistic;
    if (GRAD) S.score[0] = S.score[1] = S.score[2] = S.score[3] = 0.f;
    bool ng;
    int gi = R.gen0 + (int)alias_pick(P.gen_sel + R.gsel0, (uint32_t)R.ngen, S.rng, ng);
    const RtaGen g = P.gens[gi];
    double u = S.rng.uniform();
    S.x = g.p1x + u * g.ndx;
    S.y = g.p1y + u * g.ndy;
    if constexpr (POLY) {
        if (!poly_inside(P.sides, P.srange[g.idi], S.x, S.y)) {
            const double2 c = P.center[g.idi];
            double dx = c.x - S.x, dy = c.y - S.y;
            double dn = sqrt(dx * dx + dy * dy);
            S.x += 1.0e-6 * dx / dn;
            S.y += 1.0e-6 * dy / dn;
        }
    } else {
        const DevBox bi = P.boxes[g.idi];
        if (!box_inside(bi, S.x, S.y)) {
            double dx = 0.5 * (bi.xmin + bi.xmax) - S.x, dy = 0.5 * (bi.ymin + bi.ymax) - S.y;
            double dn = sqrt(dx * dx + dy * dy);
            S.x += 1.0e-6 * dx / dn;
            S.y += 1.0e-6 * dy / dn;
        }
    }
    S.mode = (int)alias_pick(P.gen_tab + (size_t)gi * P.n, (uint32_t)P.n, S.rng, ng);
    S.sign = g.sign;
    S.box = g.idi;
    S.t = 0.;
    S.alive = 1;
    S.need_flight = 1;
}

// Every history ends: a particle that exceeds the iteration budget or makes
// 64 events in a row without advancing in time is dropped and counted.
__device__ inline void rta_watch(const RtaParams& P, RtaState& S, RtaCounts& C) {
    S.still = (S.t == S.tlast) ? S.still + 1 : 0;
    S.tlast = S.t;
    if (++S.budget > P.max_iters || S.still > 64) {
        C.stuck++;
        S.alive = 0;
    }
}

// One step of evolparticle (RTA_MC2D.cpp:990-1386) after the flight draw of
// the history loop (2262-2288).
// POLY: the geometry has convex polygon boxes (diffuse sides only)
template <bool GRAD, int TM, bool POLY>
__device__ inline void rta_event(const RtaParams& P, void* lds, TallyAcc& A, RtaState& S, RtaCounts& C) {
    const double4 md = P.modes[S.mode];
    const DevBox B = P.boxes[S.box];
    double u = 0.;
    if (S.need_flight) {
        // exponential flight, truncated at maxtime
        if (!(S.t < P.maxtime)) {
            S.alive = 0;
            return;
        }
        u = 1.0 - S.rng.uniform();
    }
    // The scattering that ends a flight inside the box draws the next two
    // numbers of the stream (alias_pick); they are taken here on a copy of
    // the generator and the alias entry is loaded while the flight and the
    // exit test are computed. The copy replaces the generator only on that
    // path, so every stream is consumed as without the look-ahead.
    Pcg32 r2 = S.rng;
    const uint32_t sk = r2.bounded((uint32_t)P.n);
    const uint32_t su = r2.next();
    const uint2 se = P.scat[(size_t)S.scat0 + sk];
    if (S.need_flight) {
        S.dtrem = -log(u) / (md.z * S.wscale);
        if (S.t + S.dtrem > P.maxtime) S.dtrem = P.maxtime - S.t + 1.0e-8;
        if (almost_equal(S.dtrem, 0.)) return;
        S.need_flight = 0;
    }
    if (B.kind == 2) {
        // periodic box: translate, keep the remaining flight (RTA_MC2D.cpp:1034-1043)
        const int nb = POLY ? translate_poly(P.boxes, P.sides, P.srange, P.cts, P.cseg, P.ptrans, S.box, S.x, S.y,
                                             md.x, md.y, S.rng)
                            : translate_rect(P.boxes, P.cts, P.cseg, P.ptrans, S.box, S.x, S.y, md.x, md.y, S.rng);
        if (nb < 0) {
            C.badtrans++;
            S.alive = 0;
            return;
        }
        S.box = nb;
        return;
    }
    if (B.kind == 1) {  // reservoir: absorbed (RTA_MC2D.cpp:1046-1053)
        C.events++;
        S.alive = 0;
        return;
    }
    const int gmask = GRAD ? P.runs[S.run].gmask : 0;
    const double vx = md.x, vy = md.y;
    double nx = S.x + S.dtrem * vx, ny = S.y + S.dtrem * vy;
    int2 sr = make_int2(0, 0);
    if constexpr (POLY) sr = P.srange[S.box];
    const bool inside = POLY ? poly_inside(P.sides, sr, nx, ny) : box_inside(B, nx, ny);
    int mask = 0;
    double ex = 0., ey = 0., tf = 0.;
    if (!inside) {
        if constexpr (POLY) {
            tf = poly_exit(P.sides, sr, S.x, S.y, vx, vy, mask, ex, ey);
        } else {
            const double2 iv = P.ivel[S.mode];
            tf = box_exit(B, S.x, S.y, vx, vy, iv.x, iv.y, mask, ex, ey);
        }
    }
    if (inside) {
        tally_segment<GRAD, TM>(P, lds, A, S.box, S.t, S.t + S.dtrem, S.sign, vx, vy, gmask, S.score, S.mode);
        S.x = nx;
        S.y = ny;
        S.t += S.dtrem;
        if (!S.ballistic) {
            // alias_pick with the draws taken ahead
            S.mode = su < se.x ? (int)sk : (int)(se.y & 0x3fffffffu);
            S.rng = r2;
        }
        C.events++;
        S.need_flight = 1;
        return;
    }
    double left = S.dtrem - tf;
    if (almost_equal(left, 0.)) left = 0.;
    tally_segment<GRAD, TM>(P, lds, A, S.box, S.t, S.t + tf, S.sign, vx, vy, gmask, S.score, S.mode);
    S.t += tf;
    int side = 0;
    const int fast = POLY ? poly_side_fast(P.sides, sr, mask, ex, ey, vx, vy, side)
                          : side_fast(B, mask, ex, ey, vx, vy, side);
    Ids4 pb{};
    int npb;
    if (fast >= 0) {
        pb.a = fast;
        npb = 1;
    } else if (fast == SIDE_VOID) {
        npb = 0;
    } else {
        npb = POLY ? poly_find_pboxes(P.sides, P.srange, P.cts, B, ex, ey, vx, vy, pb)
                   : find_pboxes(P.boxes, P.cts, B, ex, ey, vx, vy, pb);
    }
    if (npb == 0) {
        int s = fast == SIDE_VOID ? side
              : (__popc(mask) == 1 ? __ffs(mask) - 1
                 : (POLY ? poly_pick_void_side(P.sides, P.srange, P.cts, B, sr, mask, ex, ey, S.rng.next())
                         : pick_void_side(P.boxes, P.cts, B, mask, ex, ey, S.rng.next())));
        // specular with probability p (mirror partner of the mode), else
        // diffuse; a draw is made only when 0 < p < 1 (polygon runs: diffuse)
        const float p = POLY ? 0.f : P.runs[S.run].spec[s];
        bool specular = p >= 1.f;
        if (p > 0.f && p < 1.f) {
            specular = S.rng.uniform() < (double)p;
            if (GRAD && ((gmask >> s) & 1)) {
                // d/dp log P(outcome): 1/p for specular, -1/(1-p) for diffuse
                const float d = specular ? 1.f / p : -1.f / (1.f - p);
                if (s == 0) S.score[0] += d;
                else if (s == 1) S.score[1] += d;
                else if (s == 2) S.score[2] += d;
                else S.score[3] += d;
            }
        }
        if (specular) {
            const int partner = P.mirror[(s == 0 || s == 3 ? 0 : (size_t)P.n) + S.mode];
            if (partner < 0) {
                C.badmirror++;
                S.alive = 0;
                return;
            }
            S.mode = partner;
        } else {
            // Lambert re-emission (RTA_MC2D.cpp:635-678)
            const int4 li = P.lam_info[S.mode];
            const int tabk = POLY ? P.sides[sr.x + s].nid : s;
            bool ng;
            int k = (int)alias_pick(P.lam_tab + li.z + (size_t)tabk * li.y, (uint32_t)li.y, S.rng, ng);
            S.mode = P.perm[li.x + k];
        }
        // the remaining flight is discarded (RTA_MC2D.cpp:1259)
        S.x = ex;
        S.y = ey;
        C.events++;
        S.need_flight = 1;
        return;
    }
    S.x = ex;
    S.y = ey;
    if (npb == 1) {
        S.box = pb.a;
    } else {
        // several candidate boxes at a corner (RTA_MC2D.cpp:1316-1380)
        Ids4 cand{};
        int nc = 0;
        double cx = ex + 1.0e-6 * vx, cy = ey + 1.0e-6 * vy;
        for (int k = 0; k < npb && k < 4; k++) {
            const bool in = POLY ? poly_inside(P.sides, P.srange[pb.get(k)], cx, cy)
                                 : box_inside(P.boxes[pb.get(k)], cx, cy);
            if (in) cand.set(nc++, pb.get(k));
        }
        if (nc > 0) {
            S.box = cand.get((int)S.rng.bounded((uint32_t)nc));
        } else {
            S.box = pb.get((int)S.rng.bounded((uint32_t)(npb < 4 ? npb : 4)));
            double ccx, ccy;
            if constexpr (POLY) {
                ccx = P.center[S.box].x;
                ccy = P.center[S.box].y;
            } else {
                const DevBox nb = P.boxes[S.box];
                ccx = 0.5 * (nb.xmin + nb.xmax);
                ccy = 0.5 * (nb.ymin + nb.ymax);
            }
            double dx = ccx - S.x, dy = ccy - S.y;
            double dn = sqrt(dx * dx + dy * dy);
            S.x += 1.0e-6 * dx / dn;
            S.y += 1.0e-6 * dy / dn;
        }
    }
    C.crossings++;
    S.dtrem = left;
    if (S.dtrem == 0.) S.need_flight = 1;
    return;
}

// One lane follows one particle history at a time and takes the next id
// from a per-wave pool when it ends; ids of all runs form one range, so a
// batch of runs of any sizes fills the GPU in one launch.
template <bool GRAD, int TMK, bool POLY = false>
__global__ __launch_bounds__(TMK ? 768 : 512) void rta_history_kernel(RtaParams P) {
    float* lt = nullptr;
    if constexpr (TMK == 1) {
        lt = rta_lt;
        int narr = 3;
        for (int s = 0; s < 4; s++) narr += P.lt_goff[s] >= 0 ? 1 : 0;
        const uint32_t nf = (uint32_t)narr * (uint32_t)P.lrows * (uint32_t)P.nt;
        for (uint32_t j = threadIdx.x; j < nf; j += blockDim.x) lt[j] = 0.f;
        __syncthreads();
    } else if constexpr (TMK == 2) {
        lt = rta_lt;
        double* g = (double*)rta_lt;
        const uint32_t nf = 3u * (uint32_t)P.lrows * (uint32_t)P.lbins;
        for (uint32_t j = threadIdx.x; j < nf; j += blockDim.x) g[j] = 0.;
        __syncthreads();
    }
    TallyAcc A;
    RtaState S;
    S.alive = 0;
    RtaCounts C;
    // particle ids come from a per-wave pool refilled 64 at a time, so the
    // global id counter sees one atomic per 64 histories
    unsigned long long pool_lo = 0, pool_hi = 0;
    for (;;) {
        if (S.alive) rta_watch(P, S, C);
        unsigned long long pid = ~0ULL;
        unsigned long long want = __ballot(!S.alive);
        while (want) {
            if (pool_lo == pool_hi) {
                unsigned long long base = 0;
                const int leader = __ffsll((long long)want) - 1;
                if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(P.counters + 2, 64ULL);
                base = __shfl(base, leader, 64);
                pool_lo = base;
                pool_hi = base + 64;
            }
            const unsigned long long below = want & ((1ULL << (threadIdx.x & 63)) - 1ULL);
            const unsigned long long rank = (unsigned long long)__popcll(below);
            const unsigned long long take = min((unsigned long long)__popcll(want), pool_hi - pool_lo);
            const bool mine = (want >> (threadIdx.x & 63)) & 1ULL;
            if (mine && rank < take) pid = pool_lo + rank;
            pool_lo += take;
            want = __ballot(!S.alive && pid == ~0ULL);
        }
        if (!S.alive) {
            if (pid >= P.total) break;
            rta_start<GRAD, POLY>(P, S, pid);
        }
        rta_event<GRAD, TMK, POLY>(P, lt, A, S, C);
    }
    if constexpr (TMK == 1) {
        __syncthreads();  // every history of the workgroup has ended
        lt_flush<GRAD>(P, lt);
        lt_counters(P.counters, C.events, C.crossings, C.stuck, C.badmirror, C.badtrans);
    } else if constexpr (TMK == 2) {
        acc_flush<2>(P, (double*)lt, A);
        __syncthreads();  // every lane has added its last cell
        // the workgroup's block, stored whole (rta_sum_parts_kernel adds the
        // blocks of all workgroups)
        const uint32_t nf = 3u * (uint32_t)P.lrows * (uint32_t)P.lbins;
        const double* g = (const double*)lt;
        double* dst = P.part + (size_t)blockIdx.x * nf;
        for (uint32_t j = threadIdx.x; j < nf; j += blockDim.x) dst[j] = g[j];
        lt_counters(P.counters, C.events, C.crossings, C.stuck, C.badmirror, C.badtrans);
    } else {
        if constexpr (!GRAD) acc_flush<0>(P, nullptr, A);
        block_add_u64(P.counters + 0, C.events);
        block_add_u64(P.counters + 1, C.crossings);
        block_add_u64(P.counters + 3, C.stuck);
        block_add_u64(P.counters + 4, C.badmirror);
        block_add_u64(P.counters + 5, C.badtrans);
    }
}

// Tally mode 3: persistent workgroups (768 threads, one per CU) take work
// items in order; the histories of an item (ids of one run) run with the
// run's rows, as many leading time bins as fit, in an FP64 LDS block, which
// is added into the XCD's tally copy once the item's last history has
// ended. Later bins go to the copies by device atomics, as under mode 2. A
// batch whose runs each fit a block of their own thus gets LDS sums where
// one block for all runs would not hold a single time bin.
template <bool POLY>
__global__ __launch_bounds__(768) void rta_history_items_kernel(RtaParams P) {
    int* hd = reinterpret_cast<int*>(rta_lt);  // lrows, lbins, first row, item
    double* L = reinterpret_cast<double*>(rta_lt) + 2;
    __shared__ unsigned long long next_id;
    TallyAcc A;
    RtaState S;
    S.alive = 0;
    RtaCounts C;
    const uint64_t xc = xcd_id() % RTA_NCOPY;
    for (;;) {
        __syncthreads();  // the previous item's block has been read
        if (threadIdx.x == 0) {
            const unsigned long long it = atomicAdd(P.counters + 6, 1ULL);
            hd[3] = it < (unsigned long long)P.nitems ? (int)it : P.nitems;
            if (it < (unsigned long long)P.nitems) {
                const RtaItem I = P.items[it];
                hd[0] = I.lrows;
                hd[1] = I.lbins;
                hd[2] = I.lrow0;
                next_id = I.lo;
            }
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(hd[3]);
        if (it >= P.nitems) break;
        // the item's fields as uniform scalars
        const uint64_t ihi = P.items[it].hi;
        const uint32_t lbins = (uint32_t)P.items[it].lbins, lrow0 = (uint32_t)P.items[it].lrow0;
        const uint32_t lc = (uint32_t)P.items[it].lrows * lbins, nf = 3u * lc;
        for (uint32_t j = threadIdx.x; j < nf; j += blockDim.x) L[j] = 0.;
        __syncthreads();
        // ids of the item from per-wave pools of 64, taken from next_id
        unsigned long long pool_lo = 0, pool_hi = 0;
        for (;;) {
            if (S.alive) rta_watch(P, S, C);
            unsigned long long pid = ~0ULL;
            unsigned long long want = __ballot(!S.alive);
            while (want) {
                if (pool_lo == pool_hi) {
                    unsigned long long base = 0;
                    const int leader = __ffsll((long long)want) - 1;
                    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(&next_id, 64ULL);
                    base = __shfl(base, leader, 64);
                    pool_lo = base;
                    pool_hi = base + 64;
                }
                const unsigned long long below = want & ((1ULL << (threadIdx.x & 63)) - 1ULL);
                const unsigned long long rank = (unsigned long long)__popcll(below);
                const unsigned long long take = min((unsigned long long)__popcll(want), pool_hi - pool_lo);
                const bool mine = (want >> (threadIdx.x & 63)) & 1ULL;
                if (mine && rank < take) pid = pool_lo + rank;
                pool_lo += take;
                want = __ballot(!S.alive && pid == ~0ULL);
            }
            if (!S.alive) {
                if (pid >= ihi) break;
                rta_start<false, POLY>(P, S, pid);
            }
            rta_event<false, 3, POLY>(P, L, A, S, C);
        }
        acc_flush<3>(P, L, A);
        __syncthreads();  // every history of the item has ended and added its last cell
        for (uint32_t j = threadIdx.x; j < nf; j += blockDim.x) {
            const double v = L[j];
            if (v == 0.) continue;
            const uint32_t a = j / lc, r = j - a * lc, row = r / lbins;
            const uint64_t cell = P.lrow_cell[lrow0 + row] + (r - row * lbins);
            atomicAdd(P.gz + xc * 3 * P.cells + (uint64_t)a * P.cells + cell, v);
        }
    }
    block_add_u64(P.counters + 0, C.events);
    block_add_u64(P.counters + 1, C.crossings);
    block_add_u64(P.counters + 3, C.stuck);
    block_add_u64(P.counters + 4, C.badmirror);
    block_add_u64(P.counters + 5, C.badtrans);
}

// clears the tallies and counters (a kernel rather than a memset so that the
// launch sequence can be captured in a graph)
__global__ __launch_bounds__(256) void rta_clear_kernel(double* a, uint64_t n, unsigned long long* cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = 0.;
    if (blockIdx.x == 0 && threadIdx.x < 8) cnt[threadIdx.x] = 0ULL;
}

// running sum over the time bins of every (location, mode) column: the step
// response, taken before the spectral products (they commute)
__global__ __launch_bounds__(256) void rta_spec_cumsum_kernel(double* S, int nloc, int nt, int n) {
    const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (uint64_t)nloc * n) return;
    double* p = S + (col / n) * (uint64_t)nt * n + col % n;
    double acc = 0.;
    for (int t = 0; t < nt; t++) {
        acc += p[(size_t)t * n];
        p[(size_t)t * n] = acc;
    }
}

// Tally mode 2: the LDS blocks of all workgroups, part[blk][3][lrows][lbins],
// summed per cell and added into copy 0 of the tallies (cells of the first
// lbins time bins receive no device atomics, so the add is a plain store)
__global__ __launch_bounds__(256) void rta_sum_parts_kernel(double* tally, const double* part, int nblk,
                                                            uint32_t lrows, uint32_t lbins,
                                                            const uint64_t* lrow_cell, uint64_t cells) {
    const uint32_t lc = lrows * lbins, nf = 3u * lc;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nf; j += gridDim.x * blockDim.x) {
        double v = 0.;
        for (int b = 0; b < nblk; b++) v += part[(size_t)b * nf + j];
        const uint32_t a = j / lc, r = j - a * lc, row = r / lbins;
        double* dst = tally + (uint64_t)a * cells + lrow_cell[row] + (r - row * lbins);
        *dst += v;
    }
}

// a[i] = sum over the RTA_NCOPY copies a[x n + i], for i < n
__global__ __launch_bounds__(256) void rta_sum_copies_kernel(double* a, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        double v = a[i];
        for (int x = 1; x < RTA_NCOPY; x++) v += a[(uint64_t)x * n + i];
        a[i] = v;
    }
}

struct RtaBatch {
    Ctx* ctx = nullptr;
    const Modes* m = nullptr;
    std::vector<bte_rta_run_desc> desc;
    std::vector<const Geom*> geoms;
    double thickness = 1, dt = 1, maxtime = 1, sigma = 0.25;
    int ntime = 0, nruns = 0;
    std::vector<double> timegrid;
    std::vector<double> Eeff, cv;
    std::vector<int> box0, nbox;
    std::vector<uint64_t> cell0;   // first tally cell of each run
    uint64_t cells = 0, total = 0;
    bool any_grad = false;
    bool poly = false;  // some run's geometry has polygon boxes: POLY kernels
    std::vector<RtaRunDev> runs;
    DevBuf<DevSide> d_sides;
    DevBuf<int2> d_srange;
    DevBuf<double2> d_center;
    DevBuf<double4> d_modes;
    DevBuf<double2> d_ivel;
    DevBuf<int32_t> d_mirror, d_perm;
    DevBuf<uint2> d_scat, d_gen_sel, d_gen_tab, d_lam_tab;
    DevBuf<RtaGen> d_gens;
    DevBuf<int4> d_lam_info;
    DevBuf<DevBox> d_boxes;
    DevBuf<DevContact> d_cts;
    DevBuf<double4> d_cseg;
    DevBuf<double2> d_ptrans;
    DevBuf<double> d_timegrid, d_tally, d_dgz;
    DevBuf<uint64_t> d_box_cell;
    DevBuf<RtaRunDev> d_runs, d_trial_runs;
    DevBuf<unsigned long long> d_counters;
    // spectral decomposition (one run): locations, per-mode tallies, the
    // weight matrix [n][3 nomega + nq] and the products [nloc][nt][ncols]
    int nloc = 0, nomega = 0, ncols = 0;
    std::vector<int32_t> spec_boxes;
    std::vector<double> omegagrid, spec_host;
    DevBuf<int32_t> d_box_loc;
    DevBuf<double> d_spec, d_specw, d_specout;
    int block = 256, blocks = 0;
    // tally: 0 FP64 device atomics into per-XCD copies; 1 FP32 per-workgroup
    // partial sums in LDS (one 768-thread workgroup per CU), added in FP64;
    // 2 FP64 per-workgroup partial sums of the first lbins time bins in LDS
    // (one 768-thread workgroup per CU), later bins by device atomics
    int tally = 0, lt_goff[4] = {-1, -1, -1, -1};
    uint32_t lrows = 0, lbins = 0;
    size_t lt_bytes = 0, lg_bytes = 0;  // LDS of modes 1 and 2
    DevBuf<uint32_t> d_box_lrow;
    DevBuf<uint64_t> d_lrow_cell;
    DevBuf<double> d_part;  // mode 2: [blocks][3][lrows][lbins]
    // 3: per run, the first LDS row and the rows of its regular boxes; the
    // work items and the LDS bytes of the largest run block (with header)
    std::vector<int32_t> run_lrow0, run_lrows;
    std::vector<RtaItem> items;
    DevBuf<RtaItem> d_items;
    size_t l3_bytes = 0;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ~RtaBatch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
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

// Lambert tables (RTA_MC2D.cpp:556-626): for mode i the partners f are the
// modes with nonzero omega within nsigma * hypot(sigma_i, sigma_f) in omega,
// weighted for each outward side normal n of the geometry (the four sides of
// rectangles; table k of a mode for normal k) by max(-n.v_f, 0) times the normal pdf of
// omega_i - omega_f with width hypot(sigma_i, sigma_f). sigma_i is the mode
// table's per-mode broadening (base_sigma of a q-grid) raised to the floor
// exp(p25 - 1.5 (p75 - p25)) of log sigma (RTA_MC2D.cpp:566-576), or the
// constant sigma for every mode when the table has none. The partners of i
// lie in one contiguous window of the omega-sorted order (the window of the
// widest pair); pairs outside their own cut get weight 0.
static void build_lambert(const Modes& m, double sigma, const std::vector<double2>& normals, std::vector<int4>& info,
                          std::vector<uint2>& tab, std::vector<int32_t>& perm) {
    const double nsig = 2.828427124746190097603377448419396157138;
    const double rt2pi = 2.506628274631000502415765284811045253;
    int n = m.n;
    std::vector<double> bs(n, sigma);
    if (!m.sigma.empty()) {
        std::vector<double> ls(n);
        for (int i = 0; i < n; i++) ls[i] = m.sigma[i] == 0. ? std::numeric_limits<double>::lowest() : std::log(m.sigma[i]);
        std::sort(ls.begin(), ls.end());
        const double p0 = ls[25 * (size_t)n / 100], p1 = ls[75 * (size_t)n / 100];
        const double lb = std::exp(p0 - 1.5 * (p1 - p0));
        for (int i = 0; i < n; i++) bs[i] = std::max(m.sigma[i], lb);
    }
    const double bmax = *std::max_element(bs.begin(), bs.end());
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return m.omega[a] < m.omega[b]; });
    std::vector<int> pos(n);
    for (int k = 0; k < n; k++) pos[perm[k]] = k;
    std::vector<double> sw(n);
    for (int k = 0; k < n; k++) sw[k] = m.omega[perm[k]];
    info.assign(n, make_int4(0, 0, 0, 0));
    tab.clear();
    std::vector<double> pmf;
    std::vector<uint32_t> th, al;
    for (int i = 0; i < n; i++) {
        double oi = m.omega[i];
        const double cut = nsig * std::hypot(bs[i], bmax);
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
        for (const double2& nrm : normals) {
            const double nx = nrm.x, ny = nrm.y;
            for (int k = 0; k < len; k++) {
                int f = perm[lo + k];
                const double sig = std::hypot(bs[i], bs[f]);
                double d = std::fabs(oi - m.omega[f]);
                double vn = -(nx * m.vx[f] + ny * m.vy[f]);
                bool ok = !(vn < 0. || almost_equal(vn, 0.)) && !almost_equal(0., m.omega[f]) && d <= nsig * sig;
                // the pdf's 1 / (sig sqrt(2 pi)) differs between pairs only with per-mode sigma
                const double pre = m.sigma.empty() ? 1.0 : 1.0 / (sig * rt2pi);
                pmf[k] = ok ? vn * std::exp(-0.5 * d * d / (sig * sig)) * pre : 0.;
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


// mirror partners: same omega and |v| components with one velocity
// component negated (exact float match), -1 where there is none
static void build_mirrors(const Modes& m, std::vector<int32_t>& mir) {
    const int n = m.n;
    mir.assign(2 * (size_t)n, -1);
    std::map<std::tuple<double, double, double>, int> idx;
    for (int i = 0; i < n; i++) idx.emplace(std::make_tuple(m.omega[i], m.vx[i], m.vy[i]), i);
    for (int i = 0; i < n; i++) {
        auto a = idx.find(std::make_tuple(m.omega[i], -m.vx[i], m.vy[i]));
        auto b = idx.find(std::make_tuple(m.omega[i], m.vx[i], -m.vy[i]));
        if (a != idx.end()) mir[i] = a->second;
        if (b != idx.end()) mir[(size_t)n + i] = b->second;
    }
}

static RtaParams batch_params(RtaBatch* b, const RtaRunDev* runs, uint64_t total) {
    RtaParams P{};
    P.modes = b->d_modes.p;
    P.ivel = b->d_ivel.p;
    P.mirror = b->d_mirror.p;
    P.scat = b->d_scat.p;
    P.boxes = b->d_boxes.p;
    P.cts = b->d_cts.p;
    P.cseg = b->d_cseg.p;
    P.ptrans = b->d_ptrans.p;
    P.sides = b->d_sides.p;
    P.srange = b->d_srange.p;
    P.center = b->d_center.p;
    P.gens = b->d_gens.p;
    P.gen_sel = b->d_gen_sel.p;
    P.gen_tab = b->d_gen_tab.p;
    P.lam_info = b->d_lam_info.p;
    P.lam_tab = b->d_lam_tab.p;
    P.perm = b->d_perm.p;
    P.timegrid = b->d_timegrid.p;
    P.box_cell = b->d_box_cell.p;
    P.runs = runs;
    P.gz = b->d_tally.p;
    P.jx = b->d_tally.p + b->cells;
    P.jy = b->d_tally.p + 2 * b->cells;
    P.dgz = b->any_grad ? b->d_dgz.p : nullptr;
    P.box_loc = b->nloc ? b->d_box_loc.p : nullptr;
    P.spec = b->nloc ? b->d_spec.p : nullptr;
    P.counters = b->d_counters.p;
    P.total = total;
    P.cells = b->cells;
    P.max_iters = 1ULL << 26;
    P.maxtime = b->maxtime;
    P.n = b->m->n;
    P.nt = b->ntime - 1;
    P.nruns = b->nruns;
    P.box_lrow = b->d_box_lrow.p;
    P.lrow_cell = b->d_lrow_cell.p;
    P.lrows = (int32_t)b->lrows;
    for (int k = 0; k < 4; k++) P.lt_goff[k] = b->lt_goff[k];
    P.lbins = (int32_t)b->lbins;
    P.part = b->d_part.p;
    P.idt = (double)(b->ntime - 1) / b->maxtime;
    P.items = b->d_items.p;
    P.nitems = (int32_t)b->items.size();
    return P;
}

static void launch_batch(RtaBatch* b, const RtaParams& P, int block, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(rta_clear_kernel, dim3(1024), dim3(256), 0, s, b->d_tally.p, RTA_NCOPY * 3 * b->cells,
                       b->d_counters.p);
    if (b->any_grad)
        hipLaunchKernelGGL(rta_clear_kernel, dim3(1024), dim3(256), 0, s, b->d_dgz.p, RTA_NCOPY * 4 * b->cells,
                           b->d_counters.p);
    if (b->nloc)
        hipLaunchKernelGGL(rta_clear_kernel, dim3(1024), dim3(256), 0, s, b->d_spec.p,
                           (uint64_t)b->nloc * (b->ntime - 1) * b->m->n, b->d_counters.p);
    if (b->tally == 3) {
        if (b->poly)
            hipLaunchKernelGGL(rta_history_items_kernel<true>, dim3(blocks), dim3(block), b->l3_bytes, s, P);
        else
            hipLaunchKernelGGL(rta_history_items_kernel<false>, dim3(blocks), dim3(block), b->l3_bytes, s, P);
    } else if (b->tally == 2) {
        if (b->poly)
            hipLaunchKernelGGL((rta_history_kernel<false, 2, true>), dim3(blocks), dim3(block), b->lg_bytes, s, P);
        else
            hipLaunchKernelGGL((rta_history_kernel<false, 2>), dim3(blocks), dim3(block), b->lg_bytes, s, P);
        const uint32_t nf = 3u * b->lrows * b->lbins;
        hipLaunchKernelGGL(rta_sum_parts_kernel, dim3(std::min<uint32_t>((nf + 255) / 256, 4096u)), dim3(256), 0, s,
                           b->d_tally.p, b->d_part.p, blocks, b->lrows, b->lbins, b->d_lrow_cell.p, b->cells);
    } else if (b->tally == 1) {
        if (b->any_grad)
            hipLaunchKernelGGL((rta_history_kernel<true, 1>), dim3(blocks), dim3(block), b->lt_bytes, s, P);
        else if (b->poly)
            hipLaunchKernelGGL((rta_history_kernel<false, 1, true>), dim3(blocks), dim3(block), b->lt_bytes, s, P);
        else
            hipLaunchKernelGGL((rta_history_kernel<false, 1>), dim3(blocks), dim3(block), b->lt_bytes, s, P);
    } else if (b->poly) {
        hipLaunchKernelGGL((rta_history_kernel<false, 0, true>), dim3(blocks), dim3(block), 0, s, P);
    } else if (b->any_grad) {
        hipLaunchKernelGGL((rta_history_kernel<true, 0>), dim3(blocks), dim3(block), 0, s, P);
    } else {
        hipLaunchKernelGGL((rta_history_kernel<false, 0>), dim3(blocks), dim3(block), 0, s, P);
    }
    // the per-XCD copies into copy 0
    hipLaunchKernelGGL(rta_sum_copies_kernel, dim3(1024), dim3(256), 0, s, b->d_tally.p, 3 * b->cells);
    if (b->any_grad)
        hipLaunchKernelGGL(rta_sum_copies_kernel, dim3(1024), dim3(256), 0, s, b->d_dgz.p, 4 * b->cells);
}

static void batch_create(RtaBatch* b) {
    Ctx* c = b->ctx;
    const Modes* m = b->m;
    const int n = m->n;
    hipStream_t s = c->stream;
    if (!(b->dt > 0) || !(b->maxtime > 0) || !(b->thickness > 0)) throw Error(BTE_EINVAL, "bad RTA configuration");
    for (int i = 0; i < n; i++)
        if (m->w[i] == 0.) throw Error(BTE_EINVAL, "invalid scattering rate");
    // time grid (RTA_MC2D.cpp:2121-2124)
    b->ntime = (int)std::floor(b->maxtime / b->dt + 1.0e-6) + 1;
    if (b->ntime < 2) throw Error(BTE_EINVAL, "maxtime shorter than dt");
    b->maxtime = (b->ntime - 1) * b->dt;
    b->timegrid.resize(b->ntime);
    for (int i = 0; i < b->ntime; i++) b->timegrid[i] = (double)i * b->maxtime / (b->ntime - 1);
    const int nt = b->ntime - 1;
    std::vector<double4> md(n);
    for (int i = 0; i < n; i++) md[i] = make_double4(m->vx[i], m->vy[i], m->w[i], m->omega[i]);
    b->d_modes.upload(md, s);
    std::vector<double2> iv(n);
    for (int i = 0; i < n; i++)
        iv[i] = make_double2(m->vx[i] != 0. ? 1.0 / m->vx[i] : 0., m->vy[i] != 0. ? 1.0 / m->vy[i] : 0.);
    b->d_ivel.upload(iv, s);
    std::vector<int32_t> mir;
    build_mirrors(*m, mir);
    b->d_mirror.upload(mir, s);
    const double vuc = m->V * b->thickness / m->lattz;
    std::vector<DevBox> boxes;
    std::vector<DevContact> cts;
    std::vector<double4> cseg;
    std::vector<double2> ptrans;
    std::vector<uint64_t> box_cell, lrow_cell;
    std::vector<uint32_t> box_lrow;
    std::vector<RtaGen> gens;
    std::vector<uint2> gsel, gtab, scat;
    std::vector<double> pmf(n);
    std::vector<uint32_t> th(n), al(n);
    // polygon geometries: the sides of every run's boxes, their ranges and
    // centres, and the distinct outward normals of the batch (one Lambert
    // table per normal and mode); rectangles use the four axis normals
    std::vector<DevSide> sides;
    std::vector<int2> srange;
    std::vector<double2> center, normals;
    b->poly = false;
    for (const Geom* g : b->geoms) b->poly = b->poly || (g && g->poly);
    if (!b->poly)
        for (int k = 0; k < 4; k++) {
            double nx, ny;
            side_normal(k, nx, ny);
            normals.push_back(make_double2(nx, ny));
        }
    b->nruns = (int)b->desc.size();
    b->total = 0;
    b->cells = 0;
    for (int r = 0; r < b->nruns; r++) {
        const bte_rta_run_desc& D = b->desc[r];
        const Geom* g = b->geoms[r];
        if (!g || D.nparticles == 0) throw Error(BTE_EINVAL, "run " + std::to_string(r) + ": no geometry or particles");
        if (!(D.wscale > 0)) throw Error(BTE_EINVAL, "run " + std::to_string(r) + ": wscale must be positive");
        const int nb = (int)g->boxes.size();
        for (auto& bx : g->boxes)
            if (bx.kind == 0) {
                if (bx.Teq != D.T0) throw Error(BTE_EINVAL, "Tref in xml is different from provided T0");
                if (bx.Tinit != bx.Teq) throw Error(BTE_EINVAL, "Initial temperature out of reference is not allowed");
            }
        RtaRunDev R{};
        R.first = b->total;
        R.count = D.nparticles;
        R.seed = D.seed;
        R.ballistic = D.ballistic ? 1 : 0;
        R.wscale = D.wscale;
        for (int k = 0; k < 4; k++) {
            if (!(D.specularity[k] >= 0. && D.specularity[k] <= 1.))
                throw Error(BTE_EINVAL, "specularity must lie in [0, 1]");
            R.spec[k] = (float)D.specularity[k];
            if (D.grad && D.specularity[k] > 0. && D.specularity[k] < 1.) R.gmask |= 1 << k;
            if (b->poly && (D.specularity[k] > 0. || D.grad))
                throw Error(BTE_EINVAL, "specular sides and dT/dp need rectangular boxes (diffuse sides only on polygons)");
        }
        if (R.gmask) b->any_grad = true;
        const int bo = (int)boxes.size(), co = (int)cts.size();
        b->box0.push_back(bo);
        b->nbox.push_back(nb);
        b->cell0.push_back(b->cells);
        b->run_lrow0.push_back((int32_t)b->lrows);
        for (int i = 0; i < nb; i++) {
            DevBox d = g->dboxes[i];
            if (d.kind == 2 && d.tgt < 0)
                throw Error(BTE_EGEOM, "periodic box " + std::to_string(i) + " has no translation");
            if (d.tgt >= 0) d.tgt += bo;
            ptrans.push_back(g->ptrans[i]);
            d.c0 += co;
            d.c1 += co;
            for (int k = 0; k < 4; k++)
                if (d.side_nbr[k] >= 0) d.side_nbr[k] += bo;
            boxes.push_back(d);
            if (b->poly) {
                const int2 sr = g->srange[i];
                srange.push_back(make_int2((int)sides.size(), sr.y));
                for (int k = 0; k < sr.y; k++) {
                    DevSide q = g->sides[sr.x + k];
                    if (q.nbr >= 0) q.nbr += bo;
                    int id = -1;
                    for (size_t u = 0; u < normals.size() && id < 0; u++)
                        if (std::hypot(normals[u].x - q.nx, normals[u].y - q.ny) < 1.0e-9) id = (int)u;
                    if (id < 0) {
                        id = (int)normals.size();
                        normals.push_back(make_double2(q.nx, q.ny));
                    }
                    q.nid = id;
                    sides.push_back(q);
                }
                center.push_back(make_double2(g->boxes[i].cx(), g->boxes[i].cy()));
            }
            box_cell.push_back(b->cells + (uint64_t)i * nt);
            // only regular boxes tally (reservoirs absorb, periodic boxes translate)
            box_lrow.push_back(d.kind == 0 ? b->lrows++ : ~0u);
            if (d.kind == 0) lrow_cell.push_back(b->cells + (uint64_t)i * nt);
        }
        for (auto ct : g->dcontacts) {
            ct.box += bo;
            cts.push_back(ct);
        }
        cseg.insert(cseg.end(), g->cseg.begin(), g->cseg.end());
        b->cells += (uint64_t)nb * nt;
        b->run_lrows.push_back((int32_t)b->lrows - b->run_lrow0.back());
        // BE_derivative sampler at the run temperature (sampling.cpp:47-68)
        for (int i = 0; i < n; i++) pmf[i] = m->omega[i] <= 0. ? 0. : h_be_kernel(m->omega[i], D.T0) * m->w[i];
        if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty scattering pmf");
        R.scat0 = (int32_t)scat.size();
        for (int i = 0; i < n; i++) scat.push_back(make_uint2(th[i], al[i]));
        // isothermal wall generators (RTA_MC2D.cpp:296-347, 366-397)
        std::vector<double> gw;
        R.gen0 = (int32_t)gens.size();
        for (int sid = 0; sid < nb; sid++) {
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
                gw.push_back(std::fabs(jd) * hc.length() * b->thickness);
                if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty wall pmf");
                for (int i = 0; i < n; i++) gtab.push_back(make_uint2(th[i], al[i]));
                gens.push_back(RtaGen{hc.p1x, hc.p1y, hc.p2x - hc.p1x, hc.p2y - hc.p1y, hc.box + bo, sg, 0, 0});
            }
        }
        R.ngen = (int32_t)gens.size() - R.gen0;
        if (R.ngen == 0) throw Error(BTE_EINVAL, "no isothermal generators (reservoirs at a different Teq)");
        const double einfo1 = std::accumulate(gw.begin(), gw.end(), 0.);
        b->Eeff.push_back(einfo1 / (double)D.nparticles);
        b->cv.push_back(h_calc_cv(*m, D.T0));
        std::vector<uint32_t> gth(gw.size()), gal(gw.size());
        build_alias(gw.data(), (int)gw.size(), gth.data(), gal.data());
        R.gsel0 = (int32_t)gsel.size();
        for (size_t k = 0; k < gw.size(); k++) gsel.push_back(make_uint2(gth[k], gal[k]));
        b->runs.push_back(R);
        b->total += D.nparticles;
    }
    b->d_boxes.upload(boxes, s);
    if (b->poly) {
        b->d_sides.upload(sides, s);
        b->d_srange.upload(srange, s);
        b->d_center.upload(center, s);
    }
    b->d_cts.upload(cts.empty() ? std::vector<DevContact>(1) : cts, s);
    b->d_cseg.upload(cseg.empty() ? std::vector<double4>(1) : cseg, s);
    b->d_ptrans.upload(ptrans, s);
    b->d_box_cell.upload(box_cell, s);
    b->d_box_lrow.upload(box_lrow, s);
    b->d_lrow_cell.upload(lrow_cell.empty() ? std::vector<uint64_t>(1) : lrow_cell, s);
    // LDS block of the FP32 tallies: gz, jx, jy, then one array per side that
    // some run differentiates; at least the counters' reduction space
    int gunion = 0;
    for (auto& R : b->runs) gunion |= R.gmask;
    int narr = 3;
    for (int k = 0; k < 4; k++)
        if ((gunion >> k) & 1) b->lt_goff[k] = narr++ * (int)(b->lrows * (uint32_t)nt);
    b->lt_bytes = std::max<size_t>((size_t)narr * b->lrows * nt * sizeof(float), 16 * 5 * sizeof(uint64_t));
    // mode 2: as many leading time bins of every LDS row as the workgroup's
    // LDS holds in FP64 for gz, jx and jy (the early bins carry most of the
    // segments: particles enter at t = 0 and leave through the reservoirs)
    b->lbins = b->lrows ? (uint32_t)std::min<size_t>((size_t)nt, c->lds_per_block / (3 * sizeof(double) * b->lrows)) : 0;
    b->lg_bytes = std::max<size_t>((size_t)3 * b->lrows * b->lbins * sizeof(double), 16 * 5 * sizeof(uint64_t));
    b->d_gens.upload(gens, s);
    b->d_gen_sel.upload(gsel, s);
    b->d_gen_tab.upload(gtab, s);
    b->d_scat.upload(scat, s);
    b->d_runs.upload(b->runs, s);
    // Lambert tables (shared: they depend on the modes only)
    std::vector<int4> info;
    std::vector<uint2> ltab;
    std::vector<int32_t> perm;
    build_lambert(*m, b->sigma > 0 ? b->sigma : 0.25, normals, info, ltab, perm);
    b->d_lam_info.upload(info, s);
    b->d_lam_tab.upload(ltab, s);
    b->d_perm.upload(perm, s);
    b->d_timegrid.upload(b->timegrid, s);
    b->d_tally.alloc((size_t)RTA_NCOPY * 3 * b->cells);
    if (b->any_grad) b->d_dgz.alloc((size_t)RTA_NCOPY * 4 * b->cells);
    b->d_counters.alloc(8);
    b->blocks = c->cu_count * 8;
    HIPCHECK(hipEventCreate(&b->ev0));
    HIPCHECK(hipEventCreate(&b->ev1));
    HIPCHECK(hipStreamSynchronize(s));
}

// The spectral arrays of every location as one strided-batched FP64 GEMM:
// out_l [nt][ncols] = cumsum_t(S_l) [nt][n] x W [n][ncols] (row-major)
static void spectral_products(RtaBatch* b) {
    Ctx* c = b->ctx;
    hipStream_t s = c->stream;
    const int nt = b->ntime - 1, n = b->m->n;
    const uint64_t ncolsum = (uint64_t)b->nloc * n;
    hipLaunchKernelGGL(rta_spec_cumsum_kernel, dim3((unsigned)((ncolsum + 255) / 256)), dim3(256), 0, s, b->d_spec.p,
                       b->nloc, nt, n);
    HIPCHECK(hipGetLastError());
    rocblas_handle hb = (rocblas_handle)blas_handle(c);
    const double one = 1.0, zero = 0.0;
    const rocblas_status st = rocblas_dgemm_strided_batched(
        hb, rocblas_operation_none, rocblas_operation_none, b->ncols, nt, n, &one, b->d_specw.p, b->ncols, 0,
        b->d_spec.p, n, (rocblas_stride)nt * n, &zero, b->d_specout.p, b->ncols, (rocblas_stride)nt * b->ncols,
        b->nloc);
    if (st != rocblas_status_success) throw Error(BTE_EHIP, "rocblas_dgemm_strided_batched failed");
    b->spec_host.resize((size_t)b->nloc * nt * b->ncols);
    HIPCHECK(hipMemcpyAsync(b->spec_host.data(), b->d_specout.p, b->spec_host.size() * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
}

static void batch_execute(RtaBatch* b, double* T_out, double* jx_out, double* jy_out, double* dTdp,
                          double* Eeff_out, bte_rta_batch_report* rep) {
    Ctx* c = b->ctx;
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int nt = b->ntime - 1;
    const uint64_t cells = b->cells;
    const RtaParams P = batch_params(b, b->d_runs.p, b->total);
    if (b->use_graph) {
        if (!b->gexec) {
            HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_batch(b, P, b->block, b->blocks, s);
            HIPCHECK(hipStreamEndCapture(s, &b->graph));
            HIPCHECK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        HIPCHECK(hipEventRecord(b->ev0, s));
        HIPCHECK(hipGraphLaunch(b->gexec, s));
        HIPCHECK(hipEventRecord(b->ev1, s));
    } else {
        HIPCHECK(hipEventRecord(b->ev0, s));
        launch_batch(b, P, b->block, b->blocks, s);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(b->ev1, s));
    }
    std::vector<double> tally(3 * cells), dg(b->any_grad ? 4 * cells : 0);
    unsigned long long cnt[8];
    HIPCHECK(hipMemcpyAsync(tally.data(), b->d_tally.p, 3 * cells * sizeof(double), hipMemcpyDeviceToHost, s));
    if (b->any_grad)
        HIPCHECK(hipMemcpyAsync(dg.data(), b->d_dgz.p, 4 * cells * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(cnt, b->d_counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (cnt[3] != 0)
        throw Error(BTE_EGEOM, std::to_string(cnt[3]) + " particle histories exceeded the iteration budget");
    if (cnt[4] != 0)
        throw Error(BTE_EINVAL, "specular reflection of a mode without a mirror partner in the mode table");
    if (cnt[5] != 0) throw Error(BTE_EGEOM, "Error in translation, no box found");
    if (b->nloc) spectral_products(b);
    // post-processing per run (RTA_MC2D.cpp:2336-2435): scale, cumulative
    // sum over time (step response), T = gz / (Cv / vcorr) + T0
    const double th = b->thickness, vcorr = th / b->m->lattz;
    uint64_t out0 = 0;
    for (int r = 0; r < b->nruns; r++) {
        const Geom* g = b->geoms[r];
        const int nb = b->nbox[r];
        const double E = b->Eeff[r], cvr = b->cv[r];
        double* gz = tally.data() + b->cell0[r];
        double* jx = gz + cells;
        double* jy = jx + cells;
        for (int bx = 0; bx < nb; bx++) {
            const double vbox = g->boxes[bx].area() * th;
            double agz = 0, ajx = 0, ajy = 0;
            double ad[4] = {0, 0, 0, 0};
            for (int k = 0; k < nt; k++) {
                const size_t cidx = (size_t)bx * nt + k;
                agz += gz[cidx] * E / vbox;
                ajx += jx[cidx] * E * 1e30 / vbox;
                ajy += jy[cidx] * E * 1e30 / vbox;
                const size_t o = out0 + (size_t)k * nb + bx;
                T_out[o] = agz / (cvr / vcorr) + b->desc[r].T0;
                jx_out[o] = ajx;
                jy_out[o] = ajy;
                if (dTdp) {
                    for (int q = 0; q < 4; q++) {
                        if (b->any_grad) ad[q] += dg[(size_t)q * cells + b->cell0[r] + cidx] * E / vbox;
                        dTdp[4 * out0 + (size_t)q * nt * nb + (size_t)k * nb + bx] = ad[q] / (cvr / vcorr);
                    }
                }
            }
        }
        if (Eeff_out) Eeff_out[r] = E;
        out0 += (uint64_t)nt * nb;
    }
    if (rep) {
        rep->events = cnt[0];
        rep->crossings = cnt[1];
        rep->kernel_ms = ms;
        rep->ntime = b->ntime;
        rep->block = b->block;
        rep->blocks = b->blocks;
    }
}

// Launch-shape autotuner: times the batch with every run scaled to a
// fraction of its particles for each candidate (block size, workgroups per
// CU) and keeps the fastest.
// FP32 LDS tallies on (mode 1) or off (mode 0); see bte_rta_batch_set_tally.
static void batch_set_tally(RtaBatch* b, int mode) {
    if (mode < 0 || mode > 3)
        throw Error(BTE_EINVAL,
                    "tally mode must be 0 (FP64 atomics), 1 (FP32 LDS), 2 (FP64 LDS, leading time bins) or 3 "
                    "(FP64 LDS per run, leading time bins)");
    if (mode == 3) {
        if (b->nloc) throw Error(BTE_EINVAL, "LDS tallies do not cover spectral locations");
        if (b->any_grad) throw Error(BTE_EINVAL, "tally mode 3 does not carry dT/dp; use mode 1 or 0");
        const int nt = b->ntime - 1;
        // the block: a 16-byte header, then [3][rows][bins] doubles; a margin
        // for the kernel's static LDS
        const size_t avail = b->ctx->lds_per_block - 1024;
        b->l3_bytes = 16 * 5 * sizeof(uint64_t);
        std::vector<int32_t> lb(b->nruns, 0);
        for (int r = 0; r < b->nruns; r++) {
            const size_t rows = (size_t)b->run_lrows[r];
            if (rows == 0) continue;
            if (16 + 3 * sizeof(double) * rows > avail)
                throw Error(BTE_EINVAL, "run " + std::to_string(r) + ": FP64 LDS tallies need " +
                                            std::to_string(16 + 3 * sizeof(double) * rows) +
                                            " bytes of LDS for one time bin, the device has " +
                                            std::to_string(avail) + " per workgroup");
            lb[r] = (int32_t)std::min<size_t>((size_t)nt, (avail - 16) / (3 * sizeof(double) * rows));
            b->l3_bytes = std::max(b->l3_bytes, 16 + 3 * sizeof(double) * rows * (size_t)lb[r]);
        }
        HIPCHECK(hipFuncSetAttribute(b->poly ? reinterpret_cast<const void*>(&rta_history_items_kernel<true>)
                                             : reinterpret_cast<const void*>(&rta_history_items_kernel<false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->l3_bytes));
        b->block = 768;
        b->blocks = b->ctx->cu_count;
        // about eight items per workgroup (configuration 3 without the
        // scattering look-ahead: 1.24 s, against 1.33 s with four, 1.54 s
        // with two and 1.24 s with sixteen), at
        // least 4096 ids each; the runs with most particles first, so that
        // the last items taken are short
        const uint64_t isz = std::max<uint64_t>(4096, (b->total + 8 * (uint64_t)b->blocks - 1) / (8 * (uint64_t)b->blocks));
        std::vector<int> order(b->nruns);
        for (int r = 0; r < b->nruns; r++) order[r] = r;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->runs[x].count > b->runs[y].count; });
        b->items.clear();
        for (int r : order) {
            const RtaRunDev& R = b->runs[r];
            if (R.count == 0) continue;
            const uint64_t k = (R.count + isz - 1) / isz;
            for (uint64_t q = 0; q < k; q++) {
                RtaItem I{};
                I.lo = R.first + R.count * q / k;
                I.hi = R.first + R.count * (q + 1) / k;
                I.lrow0 = b->run_lrow0[r];
                I.lrows = b->run_lrows[r];
                I.lbins = lb[r];
                I.run = r;
                b->items.push_back(I);
            }
        }
        // the same buffer while the list does not grow: a captured graph holds its address
        if (!b->items.empty()) b->d_items.upload(b->items.data(), b->items.size(), b->ctx->stream);
    } else if (mode == 2) {
        if (b->nloc) throw Error(BTE_EINVAL, "LDS tallies do not cover spectral locations");
        if (b->any_grad) throw Error(BTE_EINVAL, "tally mode 2 does not carry dT/dp; use mode 1 or 0");
        if (b->lbins == 0)
            throw Error(BTE_EINVAL, "FP64 LDS tallies need " + std::to_string(3 * sizeof(double) * b->lrows) +
                                        " bytes of LDS for one time bin, the device has " +
                                        std::to_string(b->ctx->lds_per_block) + " per workgroup");
        HIPCHECK(hipFuncSetAttribute(b->poly ? reinterpret_cast<const void*>(&rta_history_kernel<false, 2, true>)
                                             : reinterpret_cast<const void*>(&rta_history_kernel<false, 2>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->ctx->lds_per_block));
        // two 768-thread workgroups per CU when two blocks fit its LDS (small
        // geometries), else one
        b->block = 768;
        b->blocks = b->ctx->cu_count * (2 * b->lg_bytes <= b->ctx->lds_per_block ? 2 : 1);
        const size_t np = (size_t)b->blocks * 3 * b->lrows * b->lbins;
        if (b->d_part.n < np) b->d_part.alloc(np);  // kept: a captured graph holds its address
    } else if (mode == 1) {
        if (b->nloc) throw Error(BTE_EINVAL, "FP32 LDS tallies do not cover spectral locations");
        if (b->lt_bytes > b->ctx->lds_per_block)
            throw Error(BTE_EINVAL, "FP32 LDS tallies need " + std::to_string(b->lt_bytes) + " bytes of LDS, the device has " +
                                        std::to_string(b->ctx->lds_per_block) + " per workgroup");
        const void* k = b->any_grad ? reinterpret_cast<const void*>(&rta_history_kernel<true, 1>)
                        : b->poly   ? reinterpret_cast<const void*>(&rta_history_kernel<false, 1, true>)
                                    : reinterpret_cast<const void*>(&rta_history_kernel<false, 1>);
        HIPCHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->ctx->lds_per_block));
        b->block = 768;
        b->blocks = b->ctx->cu_count;
    } else if (b->tally != 0) {
        b->block = 256;
        b->blocks = b->ctx->cu_count * 8;
    }
    if (mode != b->tally && b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    b->tally = mode;
}

static void batch_tune(RtaBatch* b, double fraction) {
    if (b->tally != 0) return;  // one 768-thread workgroup per CU: the LDS holds one tally block
    Ctx* c = b->ctx;
    hipStream_t s = c->stream;
    std::vector<RtaRunDev> tr = b->runs;
    uint64_t tot = 0;
    for (auto& R : tr) {
        R.count = std::max<uint64_t>(1, (uint64_t)(R.count * fraction));
        R.first = tot;
        tot += R.count;
    }
    b->d_trial_runs.upload(tr, s);
    const RtaParams P = batch_params(b, b->d_trial_runs.p, tot);
    const int cand[][2] = {{256, 4}, {256, 8}, {256, 16}, {128, 8}, {128, 16}, {128, 32}, {512, 2}, {512, 4}};
    double best = 1e300;
    int bb = b->block, bn = b->blocks;
    for (auto& cd : cand) {
        const int blocks = c->cu_count * cd[1];
        launch_batch(b, P, cd[0], blocks, s);  // warm-up
        HIPCHECK(hipEventRecord(b->ev0, s));
        launch_batch(b, P, cd[0], blocks, s);
        HIPCHECK(hipEventRecord(b->ev1, s));
        HIPCHECK(hipStreamSynchronize(s));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        if (ms < best) {
            best = ms;
            bb = cd[0];
            bn = blocks;
        }
    }
    if (bb != b->block || bn != b->blocks) {
        b->block = bb;
        b->blocks = bn;
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
        }
    }
}

// ---- the allocation-free variant of bte_rta_execute ----
//
// Its tally copies, counters and per-workgroup LDS blocks live in this
// file-scope device array instead of buffers of the handle. Every GPU the
// library runs on holds one copy until the process ends; the host reaches it
// only through the address hipGetSymbolAddress gives for the current device
// (the array's name is not a device address in host code).
#define RTA_SCRATCH_BYTES (64ull << 20)
static __device__ __align__(256) unsigned char scratch[RTA_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device, looked up once

// the scratch array's device address on device dev (current device)
static void* scratch_address(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0) throw Error(BTE_EINVAL, "bad device");
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(BTE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Post-processing of one run on the device (RTA_MC2D.cpp:2336-2435, the
// same arithmetic as batch_execute): per box the cumulative sums over time
// of gz E / vbox, jx E 1e30 / vbox, jy E 1e30 / vbox, T = sum / (Cv / vcorr)
// + T0; outputs [nt][nbox]. One thread per box.
__global__ __launch_bounds__(256) void rta_post_kernel(const double* tally, uint64_t cells, const DevBox* boxes,
                                                       int nbox, int nt, double E, double th, double cv,
                                                       double vcorr, double T0, double* T, double* JX,
                                                       double* JY, const unsigned long long* cnt,
                                                       unsigned long long* cnt_out) {
    const int bx = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 6) cnt_out[threadIdx.x] = cnt[threadIdx.x];
    if (bx >= nbox) return;
    const DevBox b = boxes[bx];
    const double vbox = (b.xmax - b.xmin) * (b.ymax - b.ymin) * th;
    const double* gz = tally + (uint64_t)bx * nt;
    const double* jx = gz + cells;
    const double* jy = jx + cells;
    double agz = 0., ajx = 0., ajy = 0.;
    for (int k = 0; k < nt; k++) {
        agz += gz[k] * E / vbox;
        ajx += jx[k] * E * 1e30 / vbox;
        ajy += jy[k] * E * 1e30 / vbox;
        const size_t o = (size_t)k * nbox + bx;
        T[o] = agz / (cv / vcorr) + T0;
        JX[o] = ajx;
        JY[o] = ajy;
    }
}

static void rta_execute_noalloc(RtaBatch* b, double* d_T, double* d_jx, double* d_jy, unsigned long long* d_counts) {
    Ctx* c = b->ctx;
    if (b->nruns != 1) throw Error(BTE_EINVAL, "bte_rta_execute_noalloc takes a single-run handle");
    if (b->nloc) throw Error(BTE_EINVAL, "bte_rta_execute_noalloc does not tally spectral locations");
    if (b->any_grad) throw Error(BTE_EINVAL, "bte_rta_execute_noalloc does not tally the gradient sides");
    if (b->poly) throw Error(BTE_EINVAL, "bte_rta_execute_noalloc takes rectangular boxes only");
    if (!d_T || !d_jx || !d_jy || !d_counts) throw Error(BTE_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(c->device));
    double* base = (double*)scratch_address(c->device);
    // the tally mode bte_rta_execute picks, without touching the handle's
    const int mode = b->lbins > 0 ? 2 : 0;
    const int block = mode == 2 ? 768 : 256, blocks = mode == 2 ? c->cu_count : c->cu_count * 8;
    const size_t ncopy = (size_t)RTA_NCOPY * 3 * b->cells;
    const size_t ncnt = 8;  // the counters, as unsigned 64-bit words
    const size_t npart = mode == 2 ? (size_t)blocks * 3 * b->lrows * b->lbins : 0;
    const size_t need = (ncopy + ncnt + npart) * sizeof(double);
    if (need > RTA_SCRATCH_BYTES)
        throw Error(BTE_EINVAL, "bte_rta_execute_noalloc needs " + std::to_string(need) +
                                    " bytes of scratch for this run's tallies; the limit is " +
                                    std::to_string(RTA_SCRATCH_BYTES));
    RtaParams P = batch_params(b, b->d_runs.p, b->total);
    P.gz = base;
    P.jx = base + b->cells;
    P.jy = base + 2 * b->cells;
    P.dgz = nullptr;
    P.counters = (unsigned long long*)(base + ncopy);
    P.part = base + ncopy + ncnt;
    hipStream_t s = c->stream;
    // the scratch keeps the previous call's contents: the tally copies and
    // the counters start at zero (a kernel, not a memset, so that a graph
    // capture records no memset node); the partial blocks are overwritten
    hipLaunchKernelGGL(rta_clear_kernel, dim3(1024), dim3(256), 0, s, base, (uint64_t)ncopy, P.counters);
    if (mode == 2) {
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&rta_history_kernel<false, 2>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_per_block));
        hipLaunchKernelGGL((rta_history_kernel<false, 2>), dim3(blocks), dim3(block), b->lg_bytes, s, P);
        const uint32_t nf = 3u * b->lrows * b->lbins;
        hipLaunchKernelGGL(rta_sum_parts_kernel, dim3(std::min<uint32_t>((nf + 255) / 256, 4096u)), dim3(256), 0, s,
                           base, P.part, blocks, b->lrows, b->lbins, b->d_lrow_cell.p, b->cells);
    } else {
        hipLaunchKernelGGL((rta_history_kernel<false, 0>), dim3(blocks), dim3(block), 0, s, P);
    }
    hipLaunchKernelGGL(rta_sum_copies_kernel, dim3(1024), dim3(256), 0, s, base, 3 * b->cells);
    const int nb = b->nbox[0], nt = b->ntime - 1;
    const double th = b->thickness;
    hipLaunchKernelGGL(rta_post_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, base, b->cells, b->d_boxes.p, nb, nt,
                       b->Eeff[0], th, b->cv[0], th / b->m->lattz, b->desc[0].T0, d_T, d_jx, d_jy, P.counters,
                       d_counts);
    HIPCHECK(hipGetLastError());
}

}  // namespace bte

using namespace bte;

extern "C" {

int bte_rta_batch_create(bte_ctx* ctx, const bte_modes* mm, int nruns, const bte_rta_run_desc* runs,
                         double thickness, double dt, double maxtime, double sigma_bounce, int use_graph,
                         bte_rta_batch** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    auto* m = reinterpret_cast<const Modes*>(mm);
    if (!c || !m || !runs || !out || nruns <= 0) throw Error(BTE_EINVAL, "bad arguments to bte_rta_batch_create");
    HIPCHECK(hipSetDevice(c->device));
    auto* b = new RtaBatch();
    std::unique_ptr<RtaBatch> guard(b);
    b->ctx = c;
    b->m = m;
    b->desc.assign(runs, runs + nruns);
    for (int r = 0; r < nruns; r++) b->geoms.push_back(reinterpret_cast<const Geom*>(runs[r].geom));
    b->thickness = thickness;
    b->dt = dt;
    b->maxtime = maxtime;
    b->sigma = sigma_bounce;
    b->use_graph = use_graph != 0;
    batch_create(b);
    *out = reinterpret_cast<bte_rta_batch*>(guard.release());
    BTE_GUARD_END
}

int bte_rta_batch_set_tally(bte_rta_batch* bb, int mode) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(bb);
    if (!b) throw Error(BTE_EINVAL, "null batch");
    HIPCHECK(hipSetDevice(b->ctx->device));
    batch_set_tally(b, mode);
    BTE_GUARD_END
}

int bte_rta_batch_tune(bte_rta_batch* bb, double fraction, int* block, int* blocks) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(bb);
    if (!b || !(fraction > 0) || fraction > 1) throw Error(BTE_EINVAL, "bad arguments to bte_rta_batch_tune");
    HIPCHECK(hipSetDevice(b->ctx->device));
    batch_tune(b, fraction);
    if (block) *block = b->block;
    if (blocks) *blocks = b->blocks;
    BTE_GUARD_END
}

int bte_rta_batch_execute(bte_rta_batch* bb, double* T_out, double* jx_out, double* jy_out, double* dTdp,
                          double* Eeff, bte_rta_batch_report* rep) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(bb);
    if (!b || !T_out || !jx_out || !jy_out) throw Error(BTE_EINVAL, "null argument");
    batch_execute(b, T_out, jx_out, jy_out, dTdp, Eeff, rep);
    BTE_GUARD_END
}

int bte_rta_batch_destroy(bte_rta_batch* bb) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<RtaBatch*>(bb);
    BTE_GUARD_END
}

// single runs: a batch of one, with the reference's diffuse walls
int bte_rta_create(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_rta_cfg* cfg, bte_rta** out) {
    BTE_GUARD_BEGIN
    if (!cfg || !out || !g) throw Error(BTE_EINVAL, "null argument");
    if (!(cfg->dt > 0) || !(cfg->maxtime > 0) || cfg->nparticles == 0 || !(cfg->thickness > 0))
        throw Error(BTE_EINVAL, "bad RTA configuration");
    bte_rta_run_desc d{};
    d.geom = g;
    d.T0 = cfg->T0;
    d.nparticles = cfg->nparticles;
    d.seed = cfg->seed;
    d.ballistic = cfg->ballistic;
    d.wscale = 1.0;
    bte_rta_batch* b = nullptr;
    int rc = bte_rta_batch_create(ctx, m, 1, &d, cfg->thickness, cfg->dt, cfg->maxtime, cfg->sigma_bounce, 0, &b);
    if (rc != BTE_OK) return rc;
    *out = reinterpret_cast<bte_rta*>(b);
    BTE_GUARD_END
}

int bte_rta_execute(bte_rta* rr, double* T_out, double* jx_out, double* jy_out, bte_rta_report* rep) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(rr);
    if (!b || !T_out || !jx_out || !jy_out) throw Error(BTE_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(b->ctx->device));
    // FP64 LDS partial sums of the leading time bins (tally mode 2: the same
    // histories and FP64 sums) whenever one time bin of every box fits;
    // device atomics with spectral locations
    const int want = (!b->nloc && !b->any_grad && b->lbins > 0) ? 2 : 0;
    if (b->tally != want) batch_set_tally(b, want);
    bte_rta_batch_report br{};
    double E = 0;
    batch_execute(b, T_out, jx_out, jy_out, nullptr, &E, &br);
    if (rep) {
        rep->events = br.events;
        rep->crossings = br.crossings;
        rep->Eeff = E;
        rep->kernel_ms = br.kernel_ms;
        rep->ntime = br.ntime;
    }
    BTE_GUARD_END
}

int bte_rta_spectral(bte_rta* rr, int nloc, const int32_t* boxes, int nomega, const double* sigma) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(rr);
    if (!b || nloc < 0 || (nloc > 0 && (!boxes || nomega < 1))) throw Error(BTE_EINVAL, "bad arguments to bte_rta_spectral");
    if (b->nruns != 1) throw Error(BTE_EINVAL, "spectral decomposition needs a single-run handle");
    Ctx* c = b->ctx;
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const Modes* m = b->m;
    const int n = m->n, nbox = b->nbox[0], nq = m->nq;
    b->nloc = 0;
    if (nloc == 0) return BTE_OK;
    std::vector<int32_t> loc((size_t)b->box0[0] + nbox, -1);
    for (int l = 0; l < nloc; l++) {
        if (boxes[l] < 0 || boxes[l] >= nbox) throw Error(BTE_EINVAL, "spectral location out of range");
        loc[(size_t)b->box0[0] + boxes[l]] = l;
    }
    // omega ticks: LinSpaced(nomega, 0, maxomega) + delta/2, delta =
    // maxomega / nomega (RTA_MC2D.cpp:2040-2045)
    double maxomega = 0;
    for (int i = 0; i < n; i++) maxomega = std::max(maxomega, m->omega[i]);
    const double delta = maxomega / nomega;
    b->omegagrid.resize(nomega);
    for (int i = 0; i < nomega; i++)
        b->omegagrid[i] = (nomega > 1 ? maxomega * i / (nomega - 1) : maxomega) + 0.5 * delta;
    const int ncols = 3 * nomega + nq;
    std::vector<double> W;
    spectral_weights(*m, sigma ? sigma : (m->sigma.empty() ? nullptr : m->sigma.data()), b->sigma, nomega,
                     b->omegagrid.data(), W);
    const int nt = b->ntime - 1;
    b->d_box_loc.upload(loc, s);
    b->d_specw.upload(W, s);
    b->d_spec.alloc((size_t)nloc * nt * n);
    b->d_specout.alloc((size_t)nloc * nt * ncols);
    HIPCHECK(hipStreamSynchronize(s));
    b->spec_boxes.assign(boxes, boxes + nloc);
    b->nloc = nloc;
    b->nomega = nomega;
    b->ncols = ncols;
    b->spec_host.clear();
    // a captured launch would miss the spectral tally
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    BTE_GUARD_END
}

int bte_rta_spectral_get(const bte_rta* rr, int k, double* omega, double* dT, double* jx, double* jy, double* fd_q) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<const RtaBatch*>(rr);
    if (!b || k < 0 || k >= b->nloc) throw Error(BTE_EINVAL, "no such spectral location");
    if (b->spec_host.empty()) throw Error(BTE_EINVAL, "bte_rta_spectral_get before bte_rta_execute");
    const int nt = b->ntime - 1, nw = b->nomega, nq = b->m->nq, nc = b->ncols;
    const RectBox& bx = b->geoms[0]->boxes[b->spec_boxes[k]];
    const double th = b->thickness, vbox = bx.area() * th, vcorr = th / b->m->lattz, E = b->Eeff[0];
    // post-processing (RTA_MC2D.cpp:2349-2405): gz_omega * Eeff / vbox /
    // (Cv / vcorr), fluxes * Eeff 1e30 / vbox, fd * Eeff vcorr NQ /
    // (vbox 1e12 hbar); the time sums are already taken
    const double sT = E / vbox / (b->cv[0] / vcorr), sJ = E * 1e30 / vbox, sF = E * vcorr * nq / (vbox * 1.0e12 * hbar);
    if (omega) std::copy(b->omegagrid.begin(), b->omegagrid.end(), omega);
    for (int t = 0; t < nt; t++) {
        const double* r = b->spec_host.data() + ((size_t)k * nt + t) * nc;
        for (int j = 0; j < nw; j++) {
            if (dT) dT[(size_t)t * nw + j] = r[j] * sT;
            if (jx) jx[(size_t)t * nw + j] = r[nw + j] * sJ;
            if (jy) jy[(size_t)t * nw + j] = r[2 * nw + j] * sJ;
        }
        if (fd_q)
            for (int q = 0; q < nq; q++) fd_q[(size_t)t * nq + q] = r[3 * nw + q] * sF;
    }
    BTE_GUARD_END
}

int bte_spectra(bte_ctx* ctx, const bte_modes* mm, const double* sigma, double sigma_default, int nomega,
                const double* grid, int nrows, const double* X, double* out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    auto* m = reinterpret_cast<const Modes*>(mm);
    if (!c || !m || !grid || !X || !out || nomega < 1 || nrows < 0) throw Error(BTE_EINVAL, "bad arguments to bte_spectra");
    if (nrows == 0) return BTE_OK;
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    std::vector<double> W;
    spectral_weights(*m, sigma ? sigma : (m->sigma.empty() ? nullptr : m->sigma.data()), sigma_default, nomega,
                     grid, W);
    const int n = m->n, ncols = 3 * nomega + m->nq;
    DevBuf<double> dW, dX, dO;
    dW.upload(W, s);
    dX.upload(X, (size_t)nrows * n, s);
    dO.alloc((size_t)nrows * ncols);
    // out [nrows][ncols] = X [nrows][n] x W [n][ncols], row-major
    rocblas_handle hb = (rocblas_handle)blas_handle(c);
    const double one = 1.0, zero = 0.0;
    if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, ncols, nrows, n, &one, dW.p, ncols, dX.p, n,
                      &zero, dO.p, ncols) != rocblas_status_success)
        throw Error(BTE_EHIP, "rocblas_dgemm failed");
    HIPCHECK(hipMemcpyAsync(out, dO.p, (size_t)nrows * ncols * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    BTE_GUARD_END
}

int bte_rta_execute_noalloc(bte_rta* rr, double* d_T, double* d_jx, double* d_jy, unsigned long long* d_counts) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<RtaBatch*>(rr);
    if (!b) throw Error(BTE_EINVAL, "null handle");
    rta_execute_noalloc(b, d_T, d_jx, d_jy, d_counts);
    BTE_GUARD_END
}

int bte_rta_destroy(bte_rta* rr) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<RtaBatch*>(rr);
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

// ---------------------------------------------------------------------------
// RTAMC2D steady branch (nonzero <gradient>, RTA_MC2D.cpp:2544-2782): a
// generation iteration. Each generation every particle makes one flight:
// a relaxation inside a box adds its sign to phi(box) (process_segment with
// evalphi, RTA_MC2D.cpp:771-826), a diffuse wall bounce carries it into the
// next generation, a reservoir absorbs it; then |phi(box)| particles of sign
// phi are re-emitted uniformly in each box from BE_derivative. gz, jx, jy
// accumulate over generations until the relative change falls below the
// <convergence> thresholds or no particle is left.
namespace bte {

struct RtaSPart {
    double x, y;
    uint32_t mode, boxsign;  // box << 1 | (sign < 0)
    uint64_t id;
};

struct RtaSteadyParams {
    const double4* modes;
    const double2* ivel;
    const uint2* scat;
    const DevBox* boxes;
    const DevContact* cts;
    const double4* cseg;
    const double2* ptrans;
    const DevSide* sides;     // polygon geometries: sides, side range and centre per box
    const int2* srange;
    const double2* center;
    const int4* lam_info;
    const uint2* lam_tab;
    const int32_t* perm;
    const RtaGen* gens;
    const uint2* gen_sel;
    const uint2* gen_tab;
    const int32_t* gbox;     // boxes with a gradient generator
    const uint2* gbox_sel;   // alias over their energies
    const uint2* gtab;       // per gradient box: [2][n] alias (0: sign +, 1: sign -)
    const int32_t* gclass;   // per gradient box: bit 0 has +, bit 1 has -
    const uint64_t* regen;   // prefix of re-emission counts [nbox + 1]
    double* acc;             // [3][nbox] gz, jx, jy over all generations
    double* phi;             // [nbox] this generation
    const RtaSPart* pin;
    RtaSPart* pout;
    unsigned long long* counters;  // [0] appended, [1] flights, [2] stuck, [3] bad translation
    uint64_t nin, seed, key;
    double pwall;            // share of wall generators in the initial energy
    unsigned long long max_iters;
    int32_t n, nbox, ngen, ngbox, lds_tally;
};

__device__ inline void steady_append(const RtaSteadyParams& P, bool want, const RtaSPart& q) {
    const unsigned long long m = __ballot(want);
    if (m == 0ULL) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(P.counters + 0, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (want) P.pout[base + __popcll(m & ((1ULL << lane) - 1ULL))] = q;
}

// get_random_point of box b (geometry_2d.hpp:332-348): x, then y in the
// bounding box, until covered by the box (polygons; a rectangle takes the
// first draw). A box that a few thousand draws miss is counted as stuck and
// gets its centre.
template <bool POLY>
__device__ inline void steady_box_point(const RtaSteadyParams& P, int b, Pcg32& rng, double& x, double& y,
                                        unsigned long long& stuck) {
    const DevBox B = P.boxes[b];
    x = B.xmin + rng.uniform() * (B.xmax - B.xmin);
    y = B.ymin + rng.uniform() * (B.ymax - B.ymin);
    if constexpr (POLY) {
        const int2 sr = P.srange[b];
        int tries = 0;
        while (!poly_inside(P.sides, sr, x, y)) {
            if (++tries > 4096) {
                stuck++;
                x = P.center[b].x;
                y = P.center[b].y;
                return;
            }
            x = B.xmin + rng.uniform() * (B.xmax - B.xmin);
            y = B.ymin + rng.uniform() * (B.ymax - B.ymin);
        }
    }
}

// initial particles (Init_generator, RTA_MC2D.cpp:436-533)
template <bool POLY>
__global__ __launch_bounds__(256) void rta_steady_init_kernel(RtaSteadyParams P, uint64_t np) {
    unsigned long long stuck = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (uint64_t)gridDim.x * blockDim.x) {
        Pcg32 rng;
        rng.seed(P.seed, i);
        RtaSPart q{0., 0., 0u, 0u, i};
        const double r = rng.uniform();
        bool ng;
        if (P.ngen > 0 && r <= P.pwall) {
            const int gi = (int)alias_pick(P.gen_sel, (uint32_t)P.ngen, rng, ng);
            const RtaGen g = P.gens[gi];
            const double u = rng.uniform();
            double x = g.p1x + u * g.ndx, y = g.p1y + u * g.ndy;
            const DevBox bi = P.boxes[g.idi];
            if (!(POLY ? poly_inside(P.sides, P.srange[g.idi], x, y) : box_inside(bi, x, y))) {
                const double ccx = POLY ? P.center[g.idi].x : 0.5 * (bi.xmin + bi.xmax);
                const double ccy = POLY ? P.center[g.idi].y : 0.5 * (bi.ymin + bi.ymax);
                double dx = ccx - x, dy = ccy - y;
                double dn = sqrt(dx * dx + dy * dy);
                x += 1.0e-6 * dx / dn;
                y += 1.0e-6 * dy / dn;
            }
            q.x = x;
            q.y = y;
            q.mode = alias_pick(P.gen_tab + (size_t)gi * P.n, (uint32_t)P.n, rng, ng);
            q.boxsign = ((uint32_t)g.idi << 1) | (g.sign < 0 ? 1u : 0u);
        } else {
            const int k = (int)alias_pick(P.gbox_sel, (uint32_t)P.ngbox, rng, ng);
            const int b = P.gbox[k];
            steady_box_point<POLY>(P, b, rng, q.x, q.y, stuck);
            // signs alternate with the id; a class without modes gives way
            const int cls = P.gclass[k];
            int neg = (i & 1) ? 0 : 1;
            if (neg && !(cls & 2)) neg = 0;
            if (!neg && !(cls & 1)) neg = 1;
            q.mode = alias_pick(P.gtab + ((size_t)k * 2 + neg) * P.n, (uint32_t)P.n, rng, ng);
            q.boxsign = ((uint32_t)b << 1) | (uint32_t)neg;
        }
        P.pout[i] = q;
    }
    if (POLY) block_add_u64(P.counters + 2, stuck);
}

// one flight per particle (steady evolparticle, RTA_MC2D.cpp:1406-1799)
template <bool POLY>
__global__ __launch_bounds__(256) void rta_steady_flight_kernel(RtaSteadyParams P) {
    extern __shared__ __align__(16) unsigned char sdyn[];
    double* sacc = (double*)sdyn;  // [4][nbox]: gz, jx, jy, phi
    const bool lds = P.lds_tally != 0;
    if (lds)
        for (int i = threadIdx.x; i < 4 * P.nbox; i += blockDim.x) sacc[i] = 0.;
    __syncthreads();
    double* tgz = lds ? sacc : P.acc;
    double* tjx = tgz + P.nbox;
    double* tjy = tjx + P.nbox;
    double* tphi = lds ? sacc + 3 * P.nbox : P.phi;
    unsigned long long flights = 0, stuck = 0, badtrans = 0;
    const uint64_t nloop = (P.nin + 63) & ~63ULL;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i - (threadIdx.x & 63) < nloop;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const bool valid = i < P.nin;
        RtaSPart p{};
        bool bounced = false;
        if (valid) {
            p = P.pin[i];
            Pcg32 rng;
            rng.seed(P.key, i);
            int mode = (int)p.mode, box = (int)(p.boxsign >> 1);
            const int sign = (p.boxsign & 1u) ? -1 : 1;
            double4 md = P.modes[mode];
            double dt = -log(1.0 - rng.uniform()) / md.z;
            flights++;
            unsigned long long it = 0;
            int still = 0;  // consecutive crossings without advance
            double x = p.x, y = p.y;
            while (!almost_equal(dt, 0.)) {
                if (++it > P.max_iters || still > 64) { stuck++; break; }
                const DevBox B = P.boxes[box];
                if (B.kind == 2) {
                    const int nb = POLY ? translate_poly(P.boxes, P.sides, P.srange, P.cts, P.cseg, P.ptrans, box, x,
                                                         y, md.x, md.y, rng)
                                        : translate_rect(P.boxes, P.cts, P.cseg, P.ptrans, box, x, y, md.x, md.y, rng);
                    if (nb < 0) { badtrans++; break; }
                    box = nb;
                    continue;
                }
                if (B.kind == 1) break;  // absorbed
                const double vx = md.x, vy = md.y;
                const double nx = x + dt * vx, ny = y + dt * vy;
                int2 sr = make_int2(0, 0);
                if constexpr (POLY) sr = P.srange[box];
                if (POLY ? poly_inside(P.sides, sr, nx, ny) : box_inside(B, nx, ny)) {
                    atomicAdd(tgz + box, sign * dt);
                    atomicAdd(tjx + box, vx * sign * dt);
                    atomicAdd(tjy + box, vy * sign * dt);
                    atomicAdd(tphi + box, (double)sign);
                    break;
                }
                int mask;
                double ex, ey;
                double tf;
                if constexpr (POLY) {
                    tf = poly_exit(P.sides, sr, x, y, vx, vy, mask, ex, ey);
                } else {
                    const double2 iv = P.ivel[mode];
                    tf = box_exit(B, x, y, vx, vy, iv.x, iv.y, mask, ex, ey);
                }
                double left = dt - tf;
                if (almost_equal(left, 0.)) left = 0.;
                still = tf == 0. ? still + 1 : 0;
                atomicAdd(tgz + box, sign * tf);
                atomicAdd(tjx + box, vx * sign * tf);
                atomicAdd(tjy + box, vy * sign * tf);
                int side = 0;
                const int fast = POLY ? poly_side_fast(P.sides, sr, mask, ex, ey, vx, vy, side)
                                      : side_fast(B, mask, ex, ey, vx, vy, side);
                Ids4 pb{};
                int npb;
                if (fast >= 0) {
                    pb.a = fast;
                    npb = 1;
                } else if (fast == SIDE_VOID) {
                    npb = 0;
                } else {
                    npb = POLY ? poly_find_pboxes(P.sides, P.srange, P.cts, B, ex, ey, vx, vy, pb)
                               : find_pboxes(P.boxes, P.cts, B, ex, ey, vx, vy, pb);
                }
                if (npb == 0) {
                    // diffuse wall: Lambert bounce, carried to the next generation
                    const int s = fast == SIDE_VOID ? side
                                : (__popc(mask) == 1 ? __ffs(mask) - 1
                                   : (POLY ? poly_pick_void_side(P.sides, P.srange, P.cts, B, sr, mask, ex, ey, rng.next())
                                           : pick_void_side(P.boxes, P.cts, B, mask, ex, ey, rng.next())));
                    const int4 li = P.lam_info[mode];
                    const int tabk = POLY ? P.sides[sr.x + s].nid : s;
                    bool ng;
                    const int k = (int)alias_pick(P.lam_tab + li.z + (size_t)tabk * li.y, (uint32_t)li.y, rng, ng);
                    mode = P.perm[li.x + k];
                    x = ex;
                    y = ey;
                    bounced = true;
                    break;
                }
                x = ex;
                y = ey;
                if (npb == 1) {
                    box = pb.a;
                } else {
                    Ids4 cand{};
                    int nc = 0;
                    const double cx = ex + 1.0e-6 * vx, cy = ey + 1.0e-6 * vy;
                    for (int k = 0; k < npb && k < 4; k++) {
                        const bool in = POLY ? poly_inside(P.sides, P.srange[pb.get(k)], cx, cy)
                                             : box_inside(P.boxes[pb.get(k)], cx, cy);
                        if (in) cand.set(nc++, pb.get(k));
                    }
                    if (nc > 0) {
                        box = cand.get((int)rng.bounded((uint32_t)nc));
                    } else {
                        box = pb.get((int)rng.bounded((uint32_t)(npb < 4 ? npb : 4)));
                        const DevBox nb = P.boxes[box];
                        if (nb.kind != 2) {
                            const double ccx = POLY ? P.center[box].x : 0.5 * (nb.xmin + nb.xmax);
                            const double ccy = POLY ? P.center[box].y : 0.5 * (nb.ymin + nb.ymax);
                            double dx = ccx - x, dy = ccy - y;
                            double dn = sqrt(dx * dx + dy * dy);
                            x += 1.0e-6 * dx / dn;
                            y += 1.0e-6 * dy / dn;
                        }
                    }
                }
                dt = left;  // a flight ending on a crossing leaves the particle out (as the reference's loop)
            }
            p.x = x;
            p.y = y;
            p.mode = (uint32_t)mode;
            p.boxsign = ((uint32_t)box << 1) | (sign < 0 ? 1u : 0u);
        }
        steady_append(P, bounced, p);
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < 4 * P.nbox; i += blockDim.x) {
            const double v = sacc[i];
            if (v != 0.) atomicAdd(i < 3 * P.nbox ? P.acc + i : P.phi + (i - 3 * P.nbox), v);
        }
    }
    block_add_u64(P.counters + 1, flights);
    block_add_u64(P.counters + 2, stuck);
    block_add_u64(P.counters + 3, badtrans);
}

// re-emission of |phi(box)| particles per box (RTA_MC2D.cpp:2637-2654)
template <bool POLY>
__global__ __launch_bounds__(256) void rta_steady_regen_kernel(RtaSteadyParams P, uint64_t out0, uint64_t total) {
    unsigned long long stuck = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (uint64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = P.nbox - 1;  // box b holds [regen[b], regen[b+1])
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (P.regen[mid] <= k) lo = mid; else hi = mid - 1;
        }
        const int b = lo;
        Pcg32 rng;
        rng.seed(P.key ^ 0x5eed5eed5eed5eedULL, k);
        RtaSPart q;
        steady_box_point<POLY>(P, b, rng, q.x, q.y, stuck);
        bool ng;
        q.mode = alias_pick(P.scat, (uint32_t)P.n, rng, ng);
        q.boxsign = ((uint32_t)b << 1) | (P.phi[b] < 0. ? 1u : 0u);
        q.id = out0 + k;
        P.pout[out0 + k] = q;
    }
    if (POLY) block_add_u64(P.counters + 2, stuck);
}

}  // namespace bte

extern "C" int bte_rta_steady_run(bte_ctx* ctx, const bte_modes* mm, const bte_geom* gg, const bte_rta_steady_cfg* cfg,
                                  double* T_out, double* jx_out, double* jy_out, bte_rta_steady_report* rep) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    auto* m = reinterpret_cast<const Modes*>(mm);
    auto* g = reinterpret_cast<const Geom*>(gg);
    if (!c || !m || !g || !cfg || !T_out || !jx_out || !jy_out) throw Error(BTE_EINVAL, "null argument");
    if (cfg->nparticles == 0 || !(cfg->thickness > 0)) throw Error(BTE_EINVAL, "bad steady configuration");
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int n = m->n, nbox = (int)g->boxes.size();
    for (auto& b : g->boxes)
        if (b.kind == 0) {
            if (b.Teq != cfg->T0) throw Error(BTE_EINVAL, "Tref in xml is different from provided T0");
            if (b.Tinit != b.Teq) throw Error(BTE_EINVAL, "Initial temperature out of reference is not allowed");
        }
    for (int i = 0; i < n; i++)
        if (m->w[i] == 0.) throw Error(BTE_EINVAL, "invalid scattering rate");
    for (int i = 0; i < nbox; i++)
        if (g->dboxes[i].kind == 2 && g->dboxes[i].tgt < 0)
            throw Error(BTE_EGEOM, "periodic box " + std::to_string(i) + " has no translation");
    const double vuc = m->V * cfg->thickness / m->lattz;
    std::vector<double> pmf(n);
    std::vector<uint32_t> th(n), al(n);
    DevBuf<double4> d_modes;
    DevBuf<double2> d_ivel;
    std::vector<double4> md(n);
    std::vector<double2> iv(n);
    for (int i = 0; i < n; i++) {
        md[i] = make_double4(m->vx[i], m->vy[i], m->w[i], m->omega[i]);
        iv[i] = make_double2(m->vx[i] != 0. ? 1.0 / m->vx[i] : 0., m->vy[i] != 0. ? 1.0 / m->vy[i] : 0.);
    }
    d_modes.upload(md, s);
    d_ivel.upload(iv, s);
    // BE_derivative at T0 (re-emission)
    for (int i = 0; i < n; i++) pmf[i] = m->omega[i] <= 0. ? 0. : h_be_kernel(m->omega[i], cfg->T0) * m->w[i];
    if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty scattering pmf");
    DevBuf<uint2> d_scat, d_gen_sel, d_gen_tab, d_gbox_sel, d_gtab, d_lam_tab;
    upload_alias(d_scat, th, al, s);
    // wall generators (RTA_MC2D.cpp:296-347, 366-397)
    std::vector<RtaGen> gens;
    std::vector<double> gw;
    std::vector<uint2> gtab;
    for (int sid = 0; sid < nbox; sid++) {
        const RectBox& sb = g->boxes[sid];
        if (sb.kind != 1) continue;
        for (auto& hc : g->contacts[sid]) {
            const RectBox& cb = g->boxes[hc.box];
            if (cb.kind != 0 || almost_equal(hc.length(), 0.) || almost_equal(sb.Teq, cb.Teq)) continue;
            double cum = wall_pmf(*m, sb.Teq, cb.Teq, hc.npx, hc.npy, pmf.data());
            int sg = (sb.Teq - cb.Teq) >= 0 ? 1 : -1;
            gw.push_back(std::fabs(sg * cum / (m->nq * vuc)) * hc.length() * cfg->thickness);
            if (!build_alias(pmf.data(), n, th.data(), al.data())) throw Error(BTE_EINVAL, "empty wall pmf");
            for (int i = 0; i < n; i++) gtab.push_back(make_uint2(th[i], al[i]));
            gens.push_back(RtaGen{hc.p1x, hc.p1y, hc.p2x - hc.p1x, hc.p2y - hc.p1y, hc.box, sg, 0, 0});
        }
    }
    double Ewall = 0;
    for (double w : gw) Ewall += w;
    // gradient generators: Nabla_T_distribution per regular box
    // (sampling.cpp:118-187; energy vol/vuc * cumulsum / nq, RTA_MC2D.cpp:402-427)
    const double gx = cfg->gradient[0], gy = cfg->gradient[1];
    std::vector<int32_t> gbox, gclass;
    std::vector<double> gE;
    std::vector<uint2> gt;
    if (!almost_equal(std::hypot(gx, gy), 0.)) {
        std::vector<double> pp(n), pm(n);
        for (int b = 0; b < nbox; b++) {
            const RectBox& B = g->boxes[b];
            if (B.kind != 0) continue;
            double acc = 0;
            for (int i = 0; i < n; i++) {
                pp[i] = pm[i] = 0.;
                if (m->omega[i] == 0.) continue;
                const double vgt = -(gx * m->vx[i] + gy * m->vy[i]);
                const double p = std::fabs(vgt * h_be_kernel(m->omega[i], B.Teq));
                acc += p;
                (vgt >= 0. ? pp : pm)[i] = p;
            }
            int cls = 0;
            for (int sgn = 0; sgn < 2; sgn++) {
                if (build_alias((sgn == 0 ? pp : pm).data(), n, th.data(), al.data())) cls |= 1 << sgn;
                for (int i = 0; i < n; i++) gt.push_back(make_uint2(th[i], al[i]));
            }
            gbox.push_back(b);
            gclass.push_back(cls);
            gE.push_back(B.area() * cfg->thickness / vuc * kB * acc / m->nq);
        }
    }
    double Egrad = 0;
    for (double w : gE) Egrad += w;
    const double Etot = Ewall + Egrad;
    if (!(Etot > 0)) throw Error(BTE_EINVAL, "no isothermal walls or gradient sources");
    const double Eeff = Etot / (double)cfg->nparticles;
    std::vector<uint32_t> sth, sal;
    DevBuf<RtaGen> d_gens;
    if (!gens.empty()) {
        sth.resize(gw.size());
        sal.resize(gw.size());
        build_alias(gw.data(), (int)gw.size(), sth.data(), sal.data());
        upload_alias(d_gen_sel, sth, sal, s);
        d_gens.upload(gens, s);
        d_gen_tab.upload(gtab, s);
    }
    DevBuf<int32_t> d_gbox, d_gclass, d_perm, d_mirror;
    if (!gbox.empty()) {
        sth.resize(gE.size());
        sal.resize(gE.size());
        build_alias(gE.data(), (int)gE.size(), sth.data(), sal.data());
        upload_alias(d_gbox_sel, sth, sal, s);
        d_gbox.upload(gbox, s);
        d_gclass.upload(gclass, s);
        d_gtab.upload(gt, s);
    }
    std::vector<int4> info;
    std::vector<uint2> ltab;
    std::vector<int32_t> perm;
    build_lambert(*m, cfg->sigma_bounce > 0 ? cfg->sigma_bounce : 0.25, g->normals, info, ltab, perm);
    DevBuf<int4> d_lam_info;
    d_lam_info.upload(info, s);
    d_lam_tab.upload(ltab, s);
    d_perm.upload(perm, s);
    const uint64_t cap = cfg->nparticles;
    DevBuf<RtaSPart> d_parts;
    d_parts.alloc(2 * cap);
    DevBuf<double> d_acc, d_phi;
    d_acc.alloc(3 * (size_t)nbox);
    d_phi.alloc(nbox);
    DevBuf<uint64_t> d_regen;
    d_regen.alloc(nbox + 1);
    DevBuf<unsigned long long> d_cnt;
    d_cnt.alloc(4);
    HIPCHECK(hipMemsetAsync(d_acc.p, 0, 3 * (size_t)nbox * sizeof(double), s));
    RtaSteadyParams P{};
    P.modes = d_modes.p;
    P.ivel = d_ivel.p;
    P.scat = d_scat.p;
    P.boxes = g->d_boxes.p;
    P.cts = g->d_contacts.p;
    P.cseg = g->d_cseg.p;
    P.ptrans = g->d_ptrans.p;
    P.sides = g->d_sides.p;
    P.srange = g->d_srange.p;
    std::vector<double2> centre(nbox);
    for (int i = 0; i < nbox; i++) centre[i] = make_double2(g->boxes[i].cx(), g->boxes[i].cy());
    DevBuf<double2> d_centre;
    d_centre.upload(centre, s);
    P.center = d_centre.p;
    P.lam_info = d_lam_info.p;
    P.lam_tab = d_lam_tab.p;
    P.perm = d_perm.p;
    P.gens = d_gens.p;
    P.gen_sel = d_gen_sel.p;
    P.gen_tab = d_gen_tab.p;
    P.gbox = d_gbox.p;
    P.gbox_sel = d_gbox_sel.p;
    P.gtab = d_gtab.p;
    P.gclass = d_gclass.p;
    P.regen = d_regen.p;
    P.acc = d_acc.p;
    P.phi = d_phi.p;
    P.counters = d_cnt.p;
    P.seed = cfg->seed;
    P.pwall = Ewall / Etot;
    P.max_iters = 1ULL << 26;
    P.n = n;
    P.nbox = nbox;
    P.ngen = (int)gens.size();
    P.ngbox = (int)gbox.size();
    P.lds_tally = nbox <= 1024 ? 1 : 0;
    const size_t lds = P.lds_tally ? 4 * (size_t)nbox * sizeof(double) : 0;
    const int blocks = c->cu_count * 8;
    RtaSPart* buf[2] = {d_parts.p, d_parts.p + cap};
    int cur = 0;
    P.pout = buf[cur];
    hipLaunchKernelGGL(g->poly ? rta_steady_init_kernel<true> : rta_steady_init_kernel<false>, dim3(blocks), dim3(256),
                       0, s, P, cap);
    HIPCHECK(hipGetLastError());
    uint64_t nin = cap;
    std::vector<double> acc(3 * (size_t)nbox, 0.), old(3 * (size_t)nbox), phi(nbox);
    std::vector<uint64_t> pre(nbox + 1, 0);
    int gen = 0;
    bool converged = false;
    unsigned long long flights = 0;
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    struct EvG {
        hipEvent_t a, b;
        ~EvG() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } evg{e0, e1};
    double kms = 0;
    const int maxgen = cfg->max_generations > 0 ? cfg->max_generations : 1000000;
    while (nin > 0 && !converged && gen < maxgen) {
        P.pin = buf[cur];
        P.pout = buf[cur ^ 1];
        P.nin = nin;
        uint64_t z = cfg->seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(gen + 1);
        P.key = mix64(z ^ 0x57ead7ULL);
        HIPCHECK(hipMemsetAsync(d_phi.p, 0, nbox * sizeof(double), s));
        HIPCHECK(hipMemsetAsync(d_cnt.p, 0, 4 * sizeof(unsigned long long), s));
        HIPCHECK(hipEventRecord(e0, s));
        hipLaunchKernelGGL(g->poly ? rta_steady_flight_kernel<true> : rta_steady_flight_kernel<false>, dim3(blocks),
                           dim3(256), lds, s, P);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(e1, s));
        unsigned long long cnt[4];
        old = acc;
        HIPCHECK(hipMemcpyAsync(acc.data(), d_acc.p, acc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(phi.data(), d_phi.p, nbox * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(cnt, d_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        kms += ms;
        if (cnt[2]) throw Error(BTE_EGEOM, "a flight exceeded the iteration budget");
        if (cnt[3]) throw Error(BTE_EGEOM, "Error in translation, no box found");
        flights += cnt[1];
        // re-emission counts (RTA_MC2D.cpp:2637-2654); pre outlives the copy
        pre[0] = 0;
        for (int b = 0; b < nbox; b++) {
            const double a = std::fabs(phi[b]);
            pre[b + 1] = pre[b] + (almost_equal(phi[b], 0.) ? 0 : (uint64_t)(a + 1.0e-9));
        }
        const uint64_t nb = cnt[0], nre = pre[nbox];
        if (nb + nre > cap) throw Error(BTE_ECAPACITY, "steady generation larger than the first");
        if (nre) {
            d_regen.upload(pre.data(), nbox + 1, s);
            hipLaunchKernelGGL(g->poly ? rta_steady_regen_kernel<true> : rta_steady_regen_kernel<false>, dim3(blocks),
                               dim3(256), 0, s, P, nb, nre);
            HIPCHECK(hipGetLastError());
        }
        // convergence (RTA_MC2D.cpp:2660-2692)
        auto rel = [&](int q) {
            double d = 0, sq = 0;
            for (int b = 0; b < nbox; b++) {
                const double o = old[(size_t)q * nbox + b], v = acc[(size_t)q * nbox + b];
                d += (o - v) * (o - v);
                sq += v * v;
            }
            return std::sqrt(d) / std::sqrt(sq);
        };
        const double rgz = rel(0), rjx = rel(1), rjy = rel(2);
        if (almost_equal(gx, 0.)) converged = rgz < cfg->eps_energy && rjy < cfg->eps_flux;
        else if (almost_equal(gy, 0.)) converged = rgz < cfg->eps_energy && rjx < cfg->eps_flux;
        else converged = rgz < cfg->eps_energy && rjx < cfg->eps_flux && rjy < cfg->eps_flux;
        nin = nb + nre;
        cur ^= 1;
        gen++;
    }
    HIPCHECK(hipStreamSynchronize(s));
    // post-processing (RTA_MC2D.cpp:2696-2728)
    const double vcorr = cfg->thickness / m->lattz, cv = h_calc_cv(*m, cfg->T0);
    for (int b = 0; b < nbox; b++) {
        const double vbox = g->boxes[b].area() * cfg->thickness;
        T_out[b] = acc[b] / vbox * Eeff / (cv / vcorr) + cfg->T0;
        jx_out[b] = acc[nbox + b] / vbox * Eeff * 1e30;
        jy_out[b] = acc[2 * nbox + b] / vbox * Eeff * 1e30;
    }
    if (rep) {
        rep->generations = gen;
        rep->flights = flights;
        rep->Eeff = Eeff;
        rep->kernel_ms = kms;
        rep->converged = converged ? 1 : 0;
    }
    BTE_GUARD_END
}
