// Device-side building blocks shared by the MC kernels: PCG32 streams,
// alias sampling, rectangle geometry and Bose-Einstein helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bte {

constexpr double kB = 1.3806488e-23;      // include/constants.hpp:37
constexpr double hbar = 1.054571726e-34;  // include/constants.hpp:47

// PCG32 (setseq_xsh_rr_64_32): one stream per particle, state in registers.
struct Pcg32 {
    uint64_t state, inc;
    __device__ __host__ inline void seed(uint64_t initstate, uint64_t initseq) {
        inc = (initseq << 1u) | 1u;
        state = (initstate + inc) * 6364136223846793005ULL + inc;
    }
    __device__ __host__ inline uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // 53-bit uniform in [0,1)
    __device__ __host__ inline double uniform() {
        uint64_t a = next() >> 5, b = next() >> 6;
        return (double)((a << 26) | b) * (1.0 / 9007199254740992.0);
    }
    // exact uniform integer in [0, bound) (multiply-shift with rejection)
    __device__ __host__ inline uint32_t bounded(uint32_t bound) {
        uint64_t m = (uint64_t)next() * bound;
        uint32_t l = (uint32_t)m;
        if (l < bound) {
            uint32_t t = (0u - bound) % bound;
            while (l < t) {
                m = (uint64_t)next() * bound;
                l = (uint32_t)m;
            }
        }
        return (uint32_t)(m >> 32);
    }
};

// Hardware id of the XCD (accelerator complex die) the wave runs on
// (HW_REG_XCC_ID, bits 3:0), for tallies kept as one copy per XCD.
__device__ inline uint32_t xcd_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xfu; }

// Adds v of every thread of the workgroup into *dst with one global atomic.
// Counters that every wave bumped at its end cost the particle kernels tens
// of microseconds per launch: thousands of same-address atomics queue at the
// memory-side atomic unit after the last particle. blockDim.x is a multiple
// of 64 (at most 1024) and every thread of the workgroup calls this.
__device__ inline void block_add_u64(unsigned long long* dst, unsigned long long v) {
    __shared__ unsigned long long part[16];
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();  // part is reused by the next call
}

// splitmix64 finaliser: derives per-(particle, step) seeds
__device__ __host__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Alias table entry: accept column k with probability thresh/2^32, else alias.
// For propagator columns bit 31 of .y flags a negative own entry, bit 30 a
// negative alias entry; the index uses the low 30 bits.
__device__ inline uint32_t alias_pick(const uint2* tab, uint32_t n, Pcg32& r, bool& neg) {
    uint32_t k = r.bounded(n);
    uint2 e = tab[k];
    bool own = r.next() < e.x;
    neg = own ? ((e.y >> 31) & 1u) : ((e.y >> 30) & 1u);
    return own ? k : (e.y & 0x3fffffffu);
}

// include/utilities.hpp:214-223
__device__ __host__ inline bool almost_equal(double a, double b, double abstol = 1e-8,
                                             double reltol = 1e-5) {
    return fabs(a - b) < fabs(abstol) + fabs(reltol * b);
}

__device__ inline double bose_einstein(double omega, double T) {
    constexpr double pf = 1e12 * hbar / kB;
    return 1. / (exp(pf * omega / T) - 1.);
}

// Rectangle box. Sides in the reference's border order for vertex order
// (xmin,ymin),(xmin,ymax),(xmax,ymax),(xmax,ymin): 0 left, 1 bottom, 2 top,
// 3 right; outward normals (-1,0),(0,-1),(0,1),(1,0).
struct DevBox {
    double xmin, xmax, ymin, ymax;
    int32_t kind;      // 0 regular, 1 reservoir, 2 periodic
    int32_t c0, c1;    // contacts [c0, c1)
    int32_t tgt;       // periodic: the box the translation leads to
    // per side: the one box sharing the whole side (>= 0), SIDE_VOID when no
    // box shares any part of it, SIDE_MIXED otherwise
    int32_t side_nbr[4];
};
#define SIDE_VOID (-1)
#define SIDE_MIXED (-2)
struct DevContact {
    int32_t box;
    int32_t pad;
    double npx, npy;   // contact border normal, pointing out of the owner
};

__device__ __host__ inline bool box_inside(const DevBox& b, double x, double y) {
    return x >= b.xmin && x <= b.xmax && y >= b.ymin && y <= b.ymax;
}
__device__ __host__ inline void side_normal(int s, double& nx, double& ny) {
    nx = s == 0 ? -1. : (s == 3 ? 1. : 0.);
    ny = s == 1 ? -1. : (s == 2 ? 1. : 0.);
}

// Exit of the ray r + t v from box b: time of the first side crossed with
// n.v > 0 (geometry_2d.cpp:160-293 for rectangles). side_mask gets bit s for
// every side reached at that time (corner ties set two bits). ivx, ivy are
// the reciprocal velocity components (unused where the component is 0).
__device__ __host__ inline double box_exit(const DevBox& b, double x, double y, double vx,
                                           double vy, double ivx, double ivy, int& side_mask,
                                           double& ex, double& ey) {
    double tx = 1e300, ty = 1e300;
    int sx = -1, sy = -1;
    if (vx > 0.) { tx = (b.xmax - x) * ivx; sx = 3; }
    else if (vx < 0.) { tx = (b.xmin - x) * ivx; sx = 0; }
    if (vy > 0.) { ty = (b.ymax - y) * ivy; sy = 2; }
    else if (vy < 0.) { ty = (b.ymin - y) * ivy; sy = 1; }
    if (tx < 0.) tx = 0.;
    if (ty < 0.) ty = 0.;
    double t;
    if (tx < ty) {
        t = tx;
        side_mask = 1 << sx;
        ex = sx == 3 ? b.xmax : b.xmin;
        ey = y + t * vy;
    } else if (ty < tx) {
        t = ty;
        side_mask = 1 << sy;
        ex = x + t * vx;
        ey = sy == 2 ? b.ymax : b.ymin;
    } else {
        t = tx;
        side_mask = (1 << sx) | (1 << sy);
        ex = sx == 3 ? b.xmax : b.xmin;
        ey = sy == 2 ? b.ymax : b.ymin;
    }
    if (almost_equal(t, 0.)) t = 0.;
    return t;
}

// Fast path of the neighbour search for an exit through a single side at a
// point strictly inside that side: the side's precomputed neighbour decides
// (the general search below would find exactly that box, or none for a void
// side). Returns the neighbour, SIDE_VOID, or SIDE_MIXED when the general
// search must run.
__device__ __host__ inline int side_fast(const DevBox& B, int mask, double ex, double ey, double vx,
                                         double vy, int& side) {
    if (__builtin_popcount(mask) != 1) return SIDE_MIXED;
    side = __builtin_ctz(mask);
    const bool xside = (side == 0 || side == 3);
    const bool interior = xside ? (ey > B.ymin && ey < B.ymax) : (ex > B.xmin && ex < B.xmax);
    const double vn = xside ? vx : vy;
    if (!interior || almost_equal(vn, 0.)) return SIDE_MIXED;
    // select chain instead of an indexed load: keeps the box in registers
    return side == 0 ? B.side_nbr[0]
                     : (side == 1 ? B.side_nbr[1] : (side == 2 ? B.side_nbr[2] : B.side_nbr[3]));
}

// Neighbour search at an exit point (RTA_MC2D.cpp:1130-1163): contacts c
// whose closed box holds rnew or rnew + 1e-6 v/|v| with n_c.v > 0 and not
// parallel. Returns the count; the first four ids are packed in ids (kept in
// registers, no indexed local array).
struct Ids4 {
    int a, b, c, d;
    __device__ __host__ inline int get(int k) const { return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d)); }
    __device__ __host__ inline void set(int k, int v) {
        if (k == 0) a = v;
        else if (k == 1) b = v;
        else if (k == 2) c = v;
        else if (k == 3) d = v;
    }
};
__device__ __host__ inline int find_pboxes(const DevBox* boxes, const DevContact* cts,
                                           const DevBox& cur, double rx, double ry, double vx,
                                           double vy, Ids4& out) {
    double vn = sqrt(vx * vx + vy * vy);
    double ex = rx + 1.0e-6 * vx / vn, ey = ry + 1.0e-6 * vy / vn;
    int n = 0;
    for (int c = cur.c0; c < cur.c1; c++) {
        const DevContact ct = cts[c];
        double d = ct.npx * vx + ct.npy * vy;
        if (!(d > 0.) || almost_equal(d, 0.)) continue;
        const DevBox nb = boxes[ct.box];
        if (box_inside(nb, rx, ry) || box_inside(nb, ex, ey)) {
            out.set(n, ct.box);
            n++;
        }
    }
    return n;
}

// Void-side choice among tied sides (RTA_MC2D.cpp:1205-1236): a side is void
// when no contact box holds the point moved 1e-4 along its normal.
__device__ __host__ inline int pick_void_side(const DevBox* boxes, const DevContact* cts,
                                              const DevBox& cur, int mask, double rx, double ry,
                                              uint32_t rnd) {
    int c0 = -1, c1 = -1, first = -1;
    for (int s = 0; s < 4; s++) {
        if (!((mask >> s) & 1)) continue;
        if (first < 0) first = s;
        double nx, ny;
        side_normal(s, nx, ny);
        double cx = rx + 1.0e-4 * nx, cy = ry + 1.0e-4 * ny;
        bool isvoid = true;
        for (int c = cur.c0; c < cur.c1; c++)
            if (box_inside(boxes[cts[c].box], cx, cy)) { isvoid = false; break; }
        if (isvoid) {
            if (c0 < 0) c0 = s;
            else if (c1 < 0) c1 = s;
        }
    }
    if (c0 < 0) return first;
    return (c1 < 0 || (rnd & 1u) == 0) ? c0 : c1;
}

// closed-segment intersection of a0->a1 with b0->b1 (parallel segments do
// not intersect); the coordinate on an axis-aligned b is snapped to it
__device__ __host__ inline bool seg_cross(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y,
                                          double b1x, double b1y, double& px, double& py) {
    const double dx = a1x - a0x, dy = a1y - a0y, ex = b1x - b0x, ey = b1y - b0y;
    const double den = dx * ey - dy * ex;
    if (den == 0.) return false;
    const double wx = b0x - a0x, wy = b0y - a0y;
    const double t = (wx * ey - wy * ex) / den;
    const double u = (wx * dy - wy * dx) / den;
    if (t < 0. || t > 1. || u < 0. || u > 1.) return false;
    px = a0x + t * dx;
    py = a0y + t * dy;
    if (ex == 0.) px = b0x;
    if (ey == 0.) py = b0y;
    return true;
}

// geometry_2d::translate (include/geometry_2d.hpp:366-494) for rectangles.
// A particle in periodic box id at (x, y) moves along 1.0005 T (T its
// translation) to the first border (left, bottom, top, right, each extended
// by 1e-6 of its length) of the target box that this segment crosses, then
// into the regular or reservoir-free neighbour of the target box that holds
// that point: through the shared border (cseg: contact segments extended by
// 1e-6), the point itself, the point + 1e-6 v/|v| or the point - 1e-6 T/|T|.
// Several candidates: those inside after + 1e-6 v, one chosen at random.
// Returns the new box, or -1 if no box is found (a geometry error).
template <class Rng>
__device__ __host__ inline int translate_rect(const DevBox* boxes, const DevContact* cts, const double4* cseg,
                                              const double2* ptrans, int id, double& x, double& y, double vx,
                                              double vy, Rng& rng) {
    const DevBox b = boxes[id];
    const double2 T = ptrans[id];
    const double Tx = 1.0005 * T.x, Ty = 1.0005 * T.y;
    const int tib = b.tgt;
    if (tib < 0) return -1;
    const DevBox tb = boxes[tib];
    const double s0x = x, s0y = y, s1x = x + Tx, s1y = y + Ty;
    double px = 0., py = 0.;
    bool found = false;
    for (int k = 0; k < 4 && !found; k++) {
        double p1x, p1y, p2x, p2y;
        if (k == 0) { p1x = tb.xmin; p1y = tb.ymin; p2x = tb.xmin; p2y = tb.ymax; }
        else if (k == 1) { p1x = tb.xmin; p1y = tb.ymin; p2x = tb.xmax; p2y = tb.ymin; }
        else if (k == 2) { p1x = tb.xmin; p1y = tb.ymax; p2x = tb.xmax; p2y = tb.ymax; }
        else { p1x = tb.xmax; p1y = tb.ymax; p2x = tb.xmax; p2y = tb.ymin; }
        const double ndx = p2x - p1x, ndy = p2y - p1y;
        found = seg_cross(s0x, s0y, s1x, s1y, p1x - 1.0e-6 * ndx, p1y - 1.0e-6 * ndy, p2x + 1.0e-6 * ndx,
                          p2y + 1.0e-6 * ndy, px, py);
    }
    if (!found) return -1;
    const double tn = sqrt(T.x * T.x + T.y * T.y), vn = sqrt(vx * vx + vy * vy);
    const double pTx = px - 1.0e-6 * T.x / tn, pTy = py - 1.0e-6 * T.y / tn;
    const double pVx = px + 1.0e-6 * vx / vn, pVy = py + 1.0e-6 * vy / vn;
    int nb = 0, first = -1, chosen = -1, nkeep = 0;
    double fx = 0., fy = 0., cx = 0., cy = 0.;
    // two passes over the contacts: count the candidates that stay inside
    // after + 1e-6 v, then take the r-th of them
    uint32_t r = 0;
    for (int pass = 0; pass < 2; pass++) {
        int seen = 0;
        for (int c = tb.c0; c < tb.c1; c++) {
            const DevContact ct = cts[c];
            const int ibc = ct.box;
            const DevBox nbx = boxes[ibc];
            if (ibc == tib || nbx.kind != 0) continue;
            const double4 sg = cseg[c];
            double ix = 0., iy = 0.;
            const bool here_i = seg_cross(s0x, s0y, s1x, s1y, sg.x, sg.y, sg.z, sg.w, ix, iy);
            const bool here_T = box_inside(nbx, px, py);
            const bool here_Tv = box_inside(nbx, pVx, pVy);
            const bool here_TT = box_inside(nbx, pTx, pTy);
            if (!(here_i || here_T || here_Tv || here_TT)) continue;
            double rx, ry;
            if (here_i) { rx = ix; ry = iy; }
            else if (here_T) { rx = px; ry = py; }
            else if (here_Tv) { rx = pVx; ry = pVy; }
            else { rx = pTx; ry = pTy; }
            if (pass == 0) {
                if (nb == 0) { first = ibc; fx = rx; fy = ry; }
                nb++;
                if (box_inside(nbx, rx + 1.0e-6 * vx, ry + 1.0e-6 * vy)) nkeep++;
            } else if (box_inside(nbx, rx + 1.0e-6 * vx, ry + 1.0e-6 * vy)) {
                if ((uint32_t)seen == r) { chosen = ibc; cx = rx; cy = ry; break; }
                seen++;
            }
        }
        if (pass == 0) {
            if (nb == 0) return -1;
            if (nb == 1) { x = fx; y = fy; return first; }
            if (nkeep == 0) return -1;
            r = rng.bounded((uint32_t)nkeep);
        }
    }
    if (chosen < 0) return -1;
    x = cx;
    y = cy;
    return chosen;
}

// ---- convex polygon boxes ----
// Geometries with a box that is not an axis-aligned rectangle run the POLY
// kernel instantiations. Every box then has a range [first, first + count)
// of sides (int2 srange per box); each side is a hull edge p1 -> p2 in
// counter-clockwise order, with its outward unit normal, the box sharing the
// whole side (SIDE_VOID / SIDE_MIXED as for rectangles) and the id of its
// normal among the geometry's distinct normals (one diffuse-bounce table per
// normal). Per-side arrays are indexed by the global side index.
#define POLY_MAXSIDES 16
struct DevSide {
    double p1x, p1y, p2x, p2y, nx, ny;
    int32_t nbr, nid;
};

// covered_by of the closed convex polygon (geometry_2d.cpp:155-158)
__device__ __host__ inline bool poly_inside(const DevSide* S, int2 r, double x, double y) {
    for (int k = r.x; k < r.x + r.y; k++) {
        const DevSide s = S[k];
        if ((x - s.p1x) * s.nx + (y - s.p1y) * s.ny > 0.) return false;
    }
    return true;
}

// get_inter_side (geometry_2d.cpp:160-293) on a convex polygon: the first
// side with n.v > 0 that the ray r + t v reaches (a side already behind the
// start counts at t = 0); mask gets bit k for every side k (local index)
// reached at that time.
__device__ __host__ inline double poly_exit(const DevSide* S, int2 r, double x, double y, double vx, double vy,
                                            int& mask, double& ex, double& ey) {
    double t = 1e300;
    mask = 0;
    for (int k = 0; k < r.y; k++) {
        const DevSide s = S[r.x + k];
        const double vn = s.nx * vx + s.ny * vy;
        if (!(vn > 0.)) continue;
        double tk = ((s.p1x - x) * s.nx + (s.p1y - y) * s.ny) / vn;
        if (tk < 0.) tk = 0.;
        if (tk < t) {
            t = tk;
            mask = 1 << k;
        } else if (tk == t) {
            mask |= 1 << k;
        }
    }
    if (mask == 0) t = 0.;  // no side ahead (v = 0): stay
    ex = x + t * vx;
    ey = y + t * vy;
    if (almost_equal(t, 0.)) t = 0.;
    return t;
}

// fast path of the neighbour search (see side_fast): a single side reached
// away from its end points decides by its precomputed neighbour
__device__ __host__ inline int poly_side_fast(const DevSide* S, int2 r, int mask, double ex, double ey, double vx,
                                              double vy, int& side) {
    if (__builtin_popcount(mask) != 1) return SIDE_MIXED;
    side = __builtin_ctz(mask);
    const DevSide s = S[r.x + side];
    const double dx = s.p2x - s.p1x, dy = s.p2y - s.p1y;
    const double u = ((ex - s.p1x) * dx + (ey - s.p1y) * dy) / (dx * dx + dy * dy);
    if (!(u > 1.0e-9 && u < 1.0 - 1.0e-9) || almost_equal(s.nx * vx + s.ny * vy, 0.)) return SIDE_MIXED;
    return s.nbr;
}

// find_pboxes on polygons (RTA_MC2D.cpp:1130-1163)
__device__ __host__ inline int poly_find_pboxes(const DevSide* S, const int2* srange, const DevContact* cts,
                                                const DevBox& cur, double rx, double ry, double vx, double vy,
                                                Ids4& out) {
    const double vn = sqrt(vx * vx + vy * vy);
    const double ex = rx + 1.0e-6 * vx / vn, ey = ry + 1.0e-6 * vy / vn;
    int n = 0;
    for (int c = cur.c0; c < cur.c1; c++) {
        const DevContact ct = cts[c];
        const double d = ct.npx * vx + ct.npy * vy;
        if (!(d > 0.) || almost_equal(d, 0.)) continue;
        const int2 nr = srange[ct.box];
        if (poly_inside(S, nr, rx, ry) || poly_inside(S, nr, ex, ey)) {
            out.set(n, ct.box);
            n++;
        }
    }
    return n;
}

// void-side choice among tied sides (RTA_MC2D.cpp:1205-1236) on polygons;
// returns the local side index
__device__ __host__ inline int poly_pick_void_side(const DevSide* S, const int2* srange, const DevContact* cts,
                                                   const DevBox& cur, int2 r, int mask, double rx, double ry,
                                                   uint32_t rnd) {
    int c0 = -1, c1 = -1, first = -1;
    for (int k = 0; k < r.y; k++) {
        if (!((mask >> k) & 1)) continue;
        if (first < 0) first = k;
        const DevSide s = S[r.x + k];
        const double cx = rx + 1.0e-4 * s.nx, cy = ry + 1.0e-4 * s.ny;
        bool isvoid = true;
        for (int c = cur.c0; c < cur.c1; c++)
            if (poly_inside(S, srange[cts[c].box], cx, cy)) { isvoid = false; break; }
        if (isvoid) {
            if (c0 < 0) c0 = k;
            else if (c1 < 0) c1 = k;
        }
    }
    if (c0 < 0) return first;
    return (c1 < 0 || (rnd & 1u) == 0) ? c0 : c1;
}

// geometry_2d::translate (include/geometry_2d.hpp:366-494) on polygons: as
// translate_rect, with the target box's sides (each extended by 1e-6 of its
// length) as its borders
template <class Rng>
__device__ __host__ inline int translate_poly(const DevBox* boxes, const DevSide* S, const int2* srange,
                                              const DevContact* cts, const double4* cseg, const double2* ptrans,
                                              int id, double& x, double& y, double vx, double vy, Rng& rng) {
    const DevBox b = boxes[id];
    const double2 T = ptrans[id];
    const double Tx = 1.0005 * T.x, Ty = 1.0005 * T.y;
    const int tib = b.tgt;
    if (tib < 0) return -1;
    const DevBox tb = boxes[tib];
    const int2 tr = srange[tib];
    const double s0x = x, s0y = y, s1x = x + Tx, s1y = y + Ty;
    double px = 0., py = 0.;
    bool found = false;
    for (int k = tr.x; k < tr.x + tr.y && !found; k++) {
        const DevSide sd = S[k];
        const double ndx = sd.p2x - sd.p1x, ndy = sd.p2y - sd.p1y;
        found = seg_cross(s0x, s0y, s1x, s1y, sd.p1x - 1.0e-6 * ndx, sd.p1y - 1.0e-6 * ndy, sd.p2x + 1.0e-6 * ndx,
                          sd.p2y + 1.0e-6 * ndy, px, py);
    }
    if (!found) return -1;
    const double tn = sqrt(T.x * T.x + T.y * T.y), vn = sqrt(vx * vx + vy * vy);
    const double pTx = px - 1.0e-6 * T.x / tn, pTy = py - 1.0e-6 * T.y / tn;
    const double pVx = px + 1.0e-6 * vx / vn, pVy = py + 1.0e-6 * vy / vn;
    int nb = 0, first = -1, chosen = -1, nkeep = 0;
    double fx = 0., fy = 0., cx = 0., cy = 0.;
    uint32_t rsel = 0;
    for (int pass = 0; pass < 2; pass++) {
        int seen = 0;
        for (int c = tb.c0; c < tb.c1; c++) {
            const DevContact ct = cts[c];
            const int ibc = ct.box;
            if (ibc == tib || boxes[ibc].kind != 0) continue;
            const int2 nr = srange[ibc];
            const double4 sg = cseg[c];
            double ix = 0., iy = 0.;
            const bool here_i = seg_cross(s0x, s0y, s1x, s1y, sg.x, sg.y, sg.z, sg.w, ix, iy);
            const bool here_T = poly_inside(S, nr, px, py);
            const bool here_Tv = poly_inside(S, nr, pVx, pVy);
            const bool here_TT = poly_inside(S, nr, pTx, pTy);
            if (!(here_i || here_T || here_Tv || here_TT)) continue;
            double rx, ry;
            if (here_i) { rx = ix; ry = iy; }
            else if (here_T) { rx = px; ry = py; }
            else if (here_Tv) { rx = pVx; ry = pVy; }
            else { rx = pTx; ry = pTy; }
            const bool keep = poly_inside(S, nr, rx + 1.0e-6 * vx, ry + 1.0e-6 * vy);
            if (pass == 0) {
                if (nb == 0) { first = ibc; fx = rx; fy = ry; }
                nb++;
                if (keep) nkeep++;
            } else if (keep) {
                if ((uint32_t)seen == rsel) { chosen = ibc; cx = rx; cy = ry; break; }
                seen++;
            }
        }
        if (pass == 0) {
            if (nb == 0) return -1;
            if (nb == 1) { x = fx; y = fy; return first; }
            if (nkeep == 0) return -1;
            rsel = rng.bounded((uint32_t)nkeep);
        }
    }
    if (chosen < 0) return -1;
    x = cx;
    y = cy;
    return chosen;
}


}  // namespace bte
