// beRTAMC2D on the GPU: time-stepped deviational MC with a dense propagator.
//
// Per step s (reference order in src/beyondRTA_MC2d.cpp:2138-2425):
//   bert_step_kernel   cancellation of step s-1 (2295-2370) for particles
//                      that reach it, advection by dt (evolparticle 708-970),
//                      Scattering (1073-1210) or ScatteringRTA (983-1062),
//                      histogram (2281-2289), compacted write-out
//   bert_source_kernel Isothermal_walls (446-585), emitted particles
//                      advected U*dt, scattered, histogrammed (only launched
//                      when reservoirs inject)
//   bert_wallcdf_kernel adiabatic_diffuse_walls (281-435), part 1: per box
//                      side the border temperature from the J(T) spline and
//                      the emission cumulative distribution
//   bert_emit_kernel   part 2: all |Eborder| re-emissions as one balanced
//                      index space: placed, advected U*dt, scattered,
//                      histogrammed
//   bert_props_kernel  calculate_prop (1535-1602) for every box, post-cancel
//                      particle count, buffer recycling
// Particle state is 32 bytes (x, y f64; mode; box|sign; 64-bit id). The
// random stream of a particle in a step is PCG32 seeded from (seed, id,
// step), so a particle's fate does not depend on where it sits in the buffer.
#include "runtime.hpp"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <set>
#include <exception>
#include <thread>

namespace bte {

struct __align__(16) BPart {
    double x, y;
    uint32_t mode;
    uint32_t boxsign;  // box << 1 | (sign < 0)
    uint64_t pid;
};
static_assert(sizeof(BPart) == 32, "particle record is 32 bytes");

struct BertParams {
    const double2* vel;      // (vx, vy) per mode
    const double2* ivel;     // (1/vx, 1/vy) per mode (0 where the component is 0)
    const double* omega;
    const double* w;
    const DevBox* boxes;
    const DevContact* cts;
    const double4* cseg;     // contact segments (periodic translation)
    const double2* ptrans;   // periodic translation per box
    const uint2* palias;     // propagator alias per slot, ncol x ncol column-major
    const int32_t* box_pslot; // propagator slot per box (nslots > 1)
    const uint2* rta_alias;  // BE_derivative alias per box-temperature slot
    const int32_t* box_rta_slot;
    const double* box_Teq;
    const double* beq;       // [distinct regular-box Teq][n] Bose-Einstein occupations at Teq
    const int32_t* box_beq;  // per box its row of beq
    const double* spl;       // spline coefficients: per table 6 x nseg
    const int32_t* side_spl; // per side spline table index, -1 none
    // every geometry: the sides (rectangles: side s of box i at 4 i + s),
    // their box, the side range, centre and area of each box; the POLY
    // kernels take the box geometry from these
    const DevSide* sides;
    const int32_t* side_box;
    const int2* srange;
    const double2* center;
    const double* box_area;
    BPart* pin;
    BPart* pout;
    uint32_t* seg;           // [2][nsegs] live count per buffer segment
    unsigned long long* H;   // [2][nbox][n] particle counts: positive | negative << 32
    double* pbox;            // [nsegs][nbox][3] per-workgroup (sum h, sum h vx, sum h vy)
    uint32_t* tickets;       // [nbox][n]
    int32_t* Eb;             // [2][nsd]
    double* props;           // [steps+1][nbox][3]
    int64_t* npart;          // [steps+1]
    int64_t* pending;        // [steps+1] signed wall counts awaiting re-emission
    double* wcdf;            // [nsd][n] emission cumulative per side, normalised to 1
    int32_t* wguide;         // [nsd][BERT_GUIDE] first cumulative index per key
    uint32_t* wcount;        // [nsd] particles to emit this step
    int32_t* wsign;          // [nsd] their sign
    uint32_t* flags;         // [0] overflow, [1] bad mode, [2] stuck, [3] failed translation
    unsigned long long* stats;  // [0] step-kernel particle-steps, [1] emitted
    uint64_t cap;
    uint32_t segcap;         // capacity of one segment (cap / nsegs)
    int32_t nsegs;           // step + emit + source workgroups
    int32_t seg_emit0, seg_src0;  // first emit and first source segment
    int32_t ncts;
    uint64_t seed;
    double dt, Eeff, thickness;
    int32_t n, ngamma, ncol, nbox, nseg;
    int32_t nsd, poly;       // sides over all boxes; polygon boxes present
    int32_t rta, max_children;
    int32_t nslots;          // propagator slots (box temperatures)
    int32_t lds_props;       // properties from per-workgroup box sums (nbox <= BERT_LDS_BOXES)
    // dynamic LDS layout of the particle kernels (byte offsets)
    int32_t lds_pre, lds_eb, lds_sbox, lds_boxes, lds_cts, lds_vel;
    int32_t cts_lds;         // contacts staged in LDS too (GL instantiations)
    int32_t eb_lds;          // wall counters in LDS (nsd <= BERT_EB_LDS)
};

// The stream of particle pid in a given step and kernel: PCG32 with
// increment 2*pid+1 started from pid*K ^ key, where key mixes the run seed,
// the step and the kernel (a uniform value, computed once per wave).
__device__ inline uint64_t stream_key(uint64_t seed, uint32_t step, uint32_t salt) {
    return mix64(seed ^ mix64(((uint64_t)step << 8) + salt + 1));
}
__device__ inline void seed_stream(Pcg32& r, uint64_t key, uint64_t pid) {
    r.seed((pid * 0x9e3779b97f4a7c15ULL) ^ key, pid);
}

// The particle kernels share one dynamic LDS block: segment prefix sums,
// wall counters, box sums and, when small enough (GL instantiations), the
// geometry tables. Offsets are set on the host.
#define BERT_LDS_BOXES 512
#define BERT_GEO_LDS_BYTES 49152
extern __shared__ __align__(16) unsigned char bert_dyn[];
template <bool GL>
__device__ inline void geo_stage(const BertParams& P, const DevBox*& boxes, const DevContact*& cts) {
    if (GL) {
        DevBox* gb = (DevBox*)(bert_dyn + P.lds_boxes);
        DevContact* gc = (DevContact*)(bert_dyn + P.lds_cts);
        for (int i = threadIdx.x; i < P.nbox; i += blockDim.x) gb[i] = P.boxes[i];
        boxes = gb;
        cts = P.cts;
        if (P.cts_lds) {
            for (int i = threadIdx.x; i < P.ncts; i += blockDim.x) gc[i] = P.cts[i];
            cts = gc;
        }
    } else {
        boxes = P.boxes;
        cts = P.cts;
    }
}

// evolparticle (beyondRTA_MC2d.cpp:708-970): ballistic advection by dt
// through box crossings; a void side adds the particle sign to that side's
// Eborder and ends the particle. Returns false if the particle ended.
// PER compiles in the periodic translation; without periodic boxes the
// kernels use the instantiation without it, which needs about 50 fewer
// registers per lane (4 resident waves per SIMD instead of 3). POLY: convex
// polygon boxes (the geometry from the side tables).
template <bool PER, bool POLY>
__device__ inline bool advect(const BertParams& P, const DevBox* boxes, const DevContact* cts, double& x,
                              double& y, int& box, int sign, double vx, double vy, double ivx, double ivy,
                              double dt, Pcg32& rng, int32_t* Eb) {
    for (int it = 0; it < 4096; it++) {
        const DevBox B = boxes[box];
        if (B.kind == 1) return false;  // reservoir: deleted
        if (B.kind == 2) {
            if (!PER) {  // the host launches PER whenever a box is periodic
                atomicOr(P.flags + 3, 1u);
                return false;
            }
            // periodic box: translate and continue the flight (beyondRTA_MC2d.cpp:729-737)
            const int nb = POLY ? translate_poly(boxes, P.sides, P.srange, cts, P.cseg, P.ptrans, box, x, y, vx, vy, rng)
                                : translate_rect(boxes, cts, P.cseg, P.ptrans, box, x, y, vx, vy, rng);
            if (nb < 0) {
                atomicOr(P.flags + 3, 1u);
                return false;
            }
            box = nb;
            continue;
        }
        double nx = x + dt * vx, ny = y + dt * vy;
        int2 sr = make_int2(4 * box, 4);
        if constexpr (POLY) sr = P.srange[box];
        if (POLY ? poly_inside(P.sides, sr, nx, ny) : box_inside(B, nx, ny)) {
            x = nx;
            y = ny;
            return true;
        }
        int mask;
        double ex, ey;
        double tf = POLY ? poly_exit(P.sides, sr, x, y, vx, vy, mask, ex, ey)
                         : box_exit(B, x, y, vx, vy, ivx, ivy, mask, ex, ey);
        double left = dt - tf;
        if (almost_equal(left, 0.)) left = 0.;
        int side = 0;
        const int fast = POLY ? poly_side_fast(P.sides, sr, mask, ex, ey, vx, vy, side)
                              : side_fast(B, mask, ex, ey, vx, vy, side);
        if (fast == SIDE_VOID) {
            atomicAdd(Eb + sr.x + side, sign);
            return false;
        }
        if (fast >= 0) {
            x = ex;
            y = ey;
            box = fast;
            dt = left;
            if (dt == 0.) return true;
            continue;
        }
        Ids4 pb{};
        int npb = POLY ? poly_find_pboxes(P.sides, P.srange, cts, B, ex, ey, vx, vy, pb)
                       : find_pboxes(boxes, cts, B, ex, ey, vx, vy, pb);
        if (npb == 0) {
            int s = __popc(mask) == 1 ? __ffs(mask) - 1
                    : (POLY ? poly_pick_void_side(P.sides, P.srange, cts, B, sr, mask, ex, ey, rng.next())
                            : pick_void_side(boxes, cts, B, mask, ex, ey, rng.next()));
            atomicAdd(Eb + sr.x + s, sign);
            return false;
        }
        x = ex;
        y = ey;
        if (npb == 1) {
            box = pb.a;
        } else {
            Ids4 cand{};
            int nc = 0;
            double tau = 1.0e-6;
            while (nc == 0 && tau >= 1.0e-15) {
                for (int k = 0; k < npb && k < 4; k++) {
                    const double tx = ex + tau * vx, ty = ey + tau * vy;
                    const bool in = POLY ? poly_inside(P.sides, P.srange[pb.get(k)], tx, ty)
                                         : box_inside(boxes[pb.get(k)], tx, ty);
                    if (in) cand.set(nc++, pb.get(k));
                }
                tau /= 10.0;
            }
            if (nc == 0) {
                const double ccx = POLY ? P.center[box].x : 0.5 * (B.xmin + B.xmax);
                const double ccy = POLY ? P.center[box].y : 0.5 * (B.ymin + B.ymax);
                double dx = ccx - ex, dy = ccy - ey;
                double dn = sqrt(dx * dx + dy * dy);
                x = ex + 1.0e-6 * dx / dn;
                y = ey + 1.0e-6 * dy / dn;
            } else {
                box = cand.get((int)rng.bounded((uint32_t)nc));
            }
        }
        dt = left;
        if (dt == 0.) return true;
    }
    atomicOr(P.flags + 2, 1u);
    return false;
}

// The propagator of the box a particle scatters in (one slot per box
// temperature, Scattering beyondRTA_MC2d.cpp:1131-1166).
// P1 instantiations are launched when there is one propagator table and no
// RTA scattering: both tests and the slot and RTA tables drop out of the
// particle loop (fewer live scalar registers, which the generic loop spills).
template <bool P1>
__device__ inline bool rta_on(const BertParams& P) { return P1 ? false : P.rta != 0; }
template <bool P1>
__device__ inline int nslots_of(const BertParams& P) { return P1 ? 1 : P.nslots; }
template <bool P1>
__device__ inline const uint2* slot_table(const BertParams& P, int box) {
    const size_t nn = (size_t)P.ncol * P.ncol;
    return nslots_of<P1>(P) > 1 ? P.palias + (size_t)P.box_pslot[box] * nn : P.palias;
}

// The first propagator draw of a particle depends only on its mode and, with
// several box temperatures, on the box it scatters in. With one propagator it
// is drawn and its alias entry loaded before advection, so the gather
// overlaps the advection arithmetic; otherwise after advection. Later draws
// of the same particle (children of negative entries) follow in stream order.
struct Pick0 {
    uint32_t k, u;
    uint2 e;
};
template <bool P1>
__device__ inline Pick0 first_pick(const BertParams& P, Pcg32& rng, int mode, int box) {
    Pick0 f{0u, 0u, make_uint2(0u, 0u)};
    const int col = mode - P.ngamma;
    if (rta_on<P1>(P) || col < 0 || col >= P.ncol) return f;
    f.k = rng.bounded((uint32_t)P.ncol);
    f.u = rng.next();
    f.e = slot_table<P1>(P, box)[(size_t)col * P.ncol + f.k];
    return f;
}

// The (box, mode) histogram keeps the positive and the negative particle
// counts in one 64-bit word, so that one atomic tallies a particle and the
// cancellation knows how many candidates of each sign a bin holds.
__device__ inline int h_net(unsigned long long v) { return (int)(uint32_t)v - (int)(uint32_t)(v >> 32); }
// one more particle of the given sign in histogram cell h: a 32-bit atomic on
// the half that changes (positive count in the low word, negative in the high)
__device__ inline void h_add(unsigned long long* h, int sign) {
    atomicAdd(reinterpret_cast<unsigned int*>(h) + (sign > 0 ? 0 : 1), 1u);
}

// A child written in this step enters the (box, mode) histogram when that
// step feeds a cancellation or a frame (H not null) and the box sums behind
// calculate_prop (sbox in LDS, not null when the boxes fit there); vt holds
// the mode velocities (in LDS in the VL step kernels).
__device__ inline void tally_child(const BertParams& P, unsigned long long* H, double* sbox, const double2* vt, int box,
                                   int nm, int ns) {
    if (H) h_add(H + (size_t)box * P.n + nm, ns);
    if (sbox) {
        const double2 v = vt[nm];
        atomicAdd(sbox + box * 3 + 0, (double)ns);
        atomicAdd(sbox + box * 3 + 1, ns * v.x);
        atomicAdd(sbox + box * 3 + 2, ns * v.y);
    }
}

// Scattering with the propagator (beyondRTA_MC2d.cpp:1133-1192) or
// ScatteringRTA (1040-1055). Pass 0 (EMIT false) counts the children and
// returns the first one in nm0/ns0; pass 1 replays the same stream to emit
// them when there is more than one. Children of a sign flip re-queue the
// parent twice (1189-1192).
template <bool EMIT, bool P1>
__device__ inline int scatter(const BertParams& P, unsigned long long* H, double* sbox, const double2* vt, Pcg32 rng,
                              const Pick0& f0, int mode,
                              int sign, int box, BPart* out, uint64_t slot, uint64_t limit, double x, double y,
                              uint64_t pid, uint32_t step, int& nm0, int& ns0) {
    if (rta_on<P1>(P)) {
        double R = rng.uniform();
        int nm = mode;
        if (R < 1.0 - exp(-P.w[mode] * P.dt)) {
            bool ng;
            int sl = P.box_rta_slot[box];
            nm = (int)alias_pick(P.rta_alias + (size_t)sl * P.n, (uint32_t)P.n, rng, ng);
        }
        if (EMIT) {
            tally_child(P, H, sbox, vt, box, nm, sign);
            if (slot < limit) {
                BPart q{x, y, (uint32_t)nm, ((uint32_t)box << 1) | (sign < 0 ? 1u : 0u), pid};
                out[slot] = q;
            }
        }
        nm0 = nm;
        ns0 = sign;
        return 1;
    }
    int col = mode - P.ngamma;
    if (col < 0 || col >= P.ncol) {
        atomicOr(P.flags + 1, 1u);
        return 0;
    }
    const uint2* tab = slot_table<P1>(P, box) + (size_t)col * P.ncol;
    int pending = 1, nchild = 0;
    while (pending > 0 && nchild < P.max_children) {
        pending--;
        bool neg;
        int k;
        if (nchild == 0) {
            // alias_pick with the draws and entry taken before advection
            const bool own = f0.u < f0.e.x;
            neg = own ? ((f0.e.y >> 31) & 1u) : ((f0.e.y >> 30) & 1u);
            k = own ? (int)f0.k : (int)(f0.e.y & 0x3fffffffu);
        } else {
            k = (int)alias_pick(tab, (uint32_t)P.ncol, rng, neg);
        }
        int nm = k + P.ngamma;
        int ns = neg ? -sign : sign;
        if (nchild == 0) {
            nm0 = nm;
            ns0 = ns;
        }
        if (EMIT) {
            tally_child(P, H, sbox, vt, box, nm, ns);
            uint64_t s = slot + nchild;
            if (s < limit) {
                uint64_t cpid = nchild == 0 ? pid : mix64(pid ^ ((uint64_t)step << 40) ^ (uint64_t)nchild);
                BPart q{x, y, (uint32_t)nm, ((uint32_t)box << 1) | (ns < 0 ? 1u : 0u), cpid};
                out[s] = q;
            }
        }
        nchild++;
        if (neg) pending += 2;
    }
    if (pending > 0) atomicOr(P.flags + 2, 2u);
    return nchild;
}

// The particle buffers are split into segments, one per writing workgroup
// (step, emission and source kernels). A workgroup appends to its own
// segment through an LDS counter and stores the final count once, so
// compaction needs no global atomics. Readers see the segments as one list
// through the prefix sum of their counts.
#define BERT_MAXSEG 2048

// inclusive prefix of the segment counts into pre[nsegs] (blockDim.x a
// multiple of 64, at most 1024; tsum holds one entry per wave): each thread
// sums a run of segments, shuffles scan the run totals within a wave and
// the wave totals go through LDS (two barriers)
__device__ inline uint32_t seg_prefix(const uint32_t* segc, int nsegs, uint32_t segcap, uint32_t* pre,
                                      uint32_t* tsum) {
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (nsegs + nt - 1) / nt;
    const int i0 = t * per, i1 = min(nsegs, i0 + per);
    uint32_t acc = 0;
    for (int i = i0; i < i1; i++) {
        acc += min(segc[i], segcap);
        pre[i] = acc;
    }
    const int lane = t & 63, w = t >> 6;
    uint32_t v = acc;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) tsum[w] = v;
    __syncthreads();
    uint32_t off = v - acc, total = 0;
    for (int k = 0; k < (nt >> 6); k++) {
        const uint32_t x = tsum[k];
        if (k < w) off += x;
        total += x;
    }
    for (int i = i0; i < i1; i++) pre[i] += off;
    __syncthreads();
    return total;
}
// first segment s with pre[s] > i
__device__ inline int seg_search(const uint32_t* pre, int nsegs, uint32_t i) {
    int l = 0, h = nsegs - 1;
    while (l < h) {
        int mid = (l + h) >> 1;
        if (pre[mid] <= i) l = mid + 1; else h = mid;
    }
    return l;
}

// wave-wide exclusive prefix sum of c; one LDS atomic on the workgroup's
// segment counter gives the wave's base. Returns this lane's slot.
__device__ inline uint64_t wave_reserve(uint32_t* lds_count, int myseg, uint32_t segcap, int c,
                                        uint32_t* flags, uint64_t& limit) {
    const int lane = threadIdx.x & 63;
    int incl, total;
    if (__ballot(c > 1) == 0ULL) {
        // every lane has 0 or 1 child: prefix from the ballot mask
        const unsigned long long m = __ballot(c > 0);
        incl = __popcll(m & ((2ULL << lane) - 1ULL));
        total = __popcll(m);
    } else {
        incl = c;
        for (int d = 1; d < 64; d <<= 1) {
            int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        total = __shfl(incl, 63, 64);
    }
    uint32_t base = 0;
    if (lane == 63 && total > 0) base = atomicAdd(lds_count, (uint32_t)total);
    base = __shfl(base, 63, 64);
    uint64_t b = (uint64_t)base + (uint64_t)(incl - c);
    if (c > 0 && b + c > segcap) atomicOr(flags, 1u);
    limit = (uint64_t)(myseg + 1) * segcap;
    return (uint64_t)myseg * segcap + b;
}
// stores the workgroup's segment count (after all its waves appended)
__device__ inline void seg_close(const BertParams& P, uint32_t* segout, int myseg, const uint32_t* lds_count) {
    __syncthreads();
    if (threadIdx.x == 0) segout[myseg] = min(*lds_count, P.segcap);
}

// count the children, reserve their slots in the workgroup's segment and
// write them; a single child (the common case) is written from pass 0's result
template <bool P1>
__device__ inline void scatter_out(const BertParams& P, unsigned long long* H, double* sbox, const double2* vt, Pcg32& rng,
                                   const Pick0& f0,
                                   bool alive, int mode, int sign, int box, double x, double y, uint64_t pid,
                                   uint32_t step, uint32_t* lds_count, int myseg) {
    int nm0 = 0, ns0 = 1;
    int c = alive ? scatter<false, P1>(P, H, sbox, vt, rng, f0, mode, sign, box, nullptr, 0, 0, 0, 0, 0, 0, nm0, ns0) : 0;
    uint64_t limit;
    uint64_t slot = wave_reserve(lds_count, myseg, P.segcap, c, P.flags, limit);
    if (c == 1) {
        tally_child(P, H, sbox, vt, box, nm0, ns0);
        if (slot < limit) {
            BPart q{x, y, (uint32_t)nm0, ((uint32_t)box << 1) | (ns0 < 0 ? 1u : 0u), pid};
            P.pout[slot] = q;
        }
    } else if (c > 1) {
        scatter<true, P1>(P, H, sbox, vt, rng, f0, mode, sign, box, P.pout, slot, limit, x, y, pid, step, nm0, ns0);
    }
}

// A keyed pseudo-random bijection of [0, c), so that the tickets below |h|
// form a uniformly spread subset. Up to 64 candidates: the rank of the
// hashed ticket among all hashed tickets (a uniform permutation up to the
// hash). Above: a 6-round Feistel network on 2*half >= log2(c) bits with
// cycle-walking (each step is a bijection; fewer than four steps on average).
__device__ inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ inline uint32_t permute_index(uint32_t t, uint32_t c, uint64_t key) {
    if (c <= 1) return 0u;
    const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    if (c <= 64) {
        const uint32_t ht = hash32(hash32(t ^ k0) ^ k1);
        uint32_t r = 0;
        for (uint32_t u = 0; u < c; u++) {
            const uint32_t hu = hash32(hash32(u ^ k0) ^ k1);
            r += (hu < ht || (hu == ht && u < t)) ? 1u : 0u;
        }
        return r;
    }
    int b = 32 - __clz(c - 1u);
    b += b & 1;
    const int half = b >> 1;
    const uint32_t hm = (1u << half) - 1u;
    uint32_t x = t;
    do {
        uint32_t L = x >> half, R = x & hm;
        for (int r = 0; r < 6; r++) {
            const uint32_t F = hash32(R ^ ((r & 1) ? k0 : k1) ^ ((uint32_t)r * 0x9e3779b9u)) & hm;
            const uint32_t nl = R;
            R = L ^ F;
            L = nl;
        }
        x = (L << half) | R;
    } while (x >= c);
    return x;
}

// Wall counters are accumulated per workgroup in LDS and flushed with one
// global atomic per touched side (the walls of a few boxes take most hits).
#define BERT_EB_LDS 2048
__device__ inline int32_t* eb_lds_begin(const BertParams& P, int32_t* Eb, int nb4) {
    if (!P.eb_lds) return Eb;
    int32_t* sEb = (int32_t*)(bert_dyn + P.lds_eb);
    for (int i = threadIdx.x; i < nb4; i += blockDim.x) sEb[i] = 0;
    return sEb;
}
__device__ inline void eb_lds_flush(const BertParams& P, int32_t* Eb, int nb4) {
    if (!P.eb_lds) return;
    const int32_t* sEb = (const int32_t*)(bert_dyn + P.lds_eb);
    __syncthreads();
    for (int i = threadIdx.x; i < nb4; i += blockDim.x) {
        int v = sEb[i];
        if (v) atomicAdd(Eb + i, v);
    }
}

// Per-workgroup box sums for calculate_prop: cleared at kernel start, stored
// to the workgroup's row of pbox at the end (no global atomics).
__device__ inline double* box_lds_begin(const BertParams& P) {
    if (!P.lds_props) return nullptr;
    double* s = (double*)(bert_dyn + P.lds_sbox);
    for (int i = threadIdx.x; i < P.nbox * 3; i += blockDim.x) s[i] = 0.;
    return s;
}
__device__ inline void box_lds_flush(const BertParams& P, int myseg) {
    if (!P.lds_props) return;
    const double* s = (const double*)(bert_dyn + P.lds_sbox);
    __syncthreads();
    double* o = P.pbox + (size_t)myseg * P.nbox * 3;
    for (int i = threadIdx.x; i < P.nbox * 3; i += blockDim.x) o[i] = s[i];
}

// One workgroup per segment of the step's output. Workgroup w takes the
// w-th equal slice of the live particle list (all input segments).
// P1 without periodic boxes fits 96 VGPRs: five workgroups (waves per SIMD)
// stay resident per CU
// VL: one 1024-thread workgroup per CU with the mode velocities and their
// inverses (32 B per mode) copied into LDS after the shared layout, so the
// per-particle table reads are LDS reads; otherwise 256-thread workgroups,
// five per CU for the P1 kernels, reading the tables from global memory.
template <bool GL, bool PER, bool P1, bool POLY, bool VL = false>
__global__ __launch_bounds__(VL ? 1024 : 256, VL ? 1 : ((P1 && !PER) ? 5 : 1)) void bert_step_kernel(
    BertParams P, uint32_t step, int cur, int cancel_prev, int hist) {
    __shared__ uint32_t tsum[16];
    __shared__ uint32_t lcount;
    uint32_t* pre = (uint32_t*)(bert_dyn + P.lds_pre);
    const DevBox* boxes;
    const DevContact* cts;
    geo_stage<GL>(P, boxes, cts);
    const double2* vtab = P.vel;
    const double2* ivtab = P.ivel;
    if constexpr (VL) {
        // visible after seg_prefix's barrier below
        double2* sv = (double2*)(bert_dyn + P.lds_vel);
        for (int q = threadIdx.x; q < P.n; q += blockDim.x) {
            sv[q] = P.vel[q];
            sv[P.n + q] = P.ivel[q];
        }
        vtab = sv;
        ivtab = sv + P.n;
    }
    const int nb4 = P.nsd;
    int32_t* Eg = P.Eb + (size_t)(step & 1) * nb4;
    int32_t* Eb = eb_lds_begin(P, Eg, nb4);
    if (threadIdx.x == 0) lcount = 0;
    double* sbox = box_lds_begin(P);
    const uint32_t* segin = P.seg + (size_t)cur * P.nsegs;
    uint32_t* segout = P.seg + (size_t)(cur ^ 1) * P.nsegs;
    const uint32_t nin = seg_prefix(segin, P.nsegs, P.segcap, pre, tsum);  // ends with a barrier
    const uint32_t lo = (uint32_t)((uint64_t)nin * blockIdx.x / gridDim.x);
    const uint32_t hi = (uint32_t)((uint64_t)nin * (blockIdx.x + 1) / gridDim.x);
    const int myseg = blockIdx.x;
    unsigned long long* H = hist ? P.H + (size_t)cur * P.nbox * P.n : nullptr;
    const unsigned long long* Hprev = P.H + (size_t)(cur ^ 1) * P.nbox * P.n;
    const uint64_t skey = stream_key(P.seed, step, 0);
    const uint64_t ckey = stream_key(P.seed, step, 5);
    unsigned long long done = 0;
    // segment of the wave's first index, advanced as the wave moves on
    int sgw = lo < hi ? seg_search(pre, P.nsegs, lo) : 0;
    for (uint32_t base = lo; base < hi; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint32_t wfirst = base + (threadIdx.x & ~63u);
        while (sgw < P.nsegs - 1 && wfirst >= pre[sgw]) sgw++;
        int sg = sgw;
        BPart p{};
        int box = 0, sign = 1, mode = 0;
        bool alive = false;
        Pcg32 rng;
        Pick0 f0{0u, 0u, make_uint2(0u, 0u)};
        if (valid) {
            while (sg < P.nsegs - 1 && i >= pre[sg]) sg++;
            const uint32_t off = i - (sg > 0 ? pre[sg - 1] : 0u);
            p = P.pin[(uint64_t)sg * P.segcap + off];
            box = (int)(p.boxsign >> 1);
            sign = (p.boxsign & 1u) ? -1 : 1;
            mode = (int)p.mode;
            alive = true;
            if (cancel_prev) {
                // cancellation (beyondRTA_MC2d.cpp:2342-2359): keep |h| particles
                // of sign(h) per (box, mode). Which ones is unspecified in the
                // reference (an unstable sort orders them): the c candidates
                // take tickets 0..c-1 and a keyed permutation of the tickets
                // picks the |h| survivors, a subset that does not depend on
                // the order in which the candidates arrive.
                const size_t hb = (size_t)box * P.n + mode;
                const unsigned long long hv = Hprev[hb];
                const int h = h_net(hv);
                if (h == 0 || ((sign < 0) != (h < 0))) {
                    alive = false;
                } else {
                    const uint32_t c = h < 0 ? (uint32_t)(hv >> 32) : (uint32_t)hv;
                    const uint32_t keep = (uint32_t)(h < 0 ? -h : h);
                    // every candidate survives when the bin holds one sign
                    if (keep < c) {
                        const uint32_t tk = atomicAdd(P.tickets + hb, 1u);
                        if (tk >= c || permute_index(tk, c, mix64(ckey ^ (uint64_t)hb)) >= keep) alive = false;
                    }
                }
            }
            if (alive) {
                seed_stream(rng, skey, p.pid);
                if (nslots_of<P1>(P) == 1) f0 = first_pick<P1>(P, rng, mode, 0);
                const double2 v = vtab[mode];
                const double2 iv = ivtab[mode];
                alive = advect<PER, POLY>(P, boxes, cts, p.x, p.y, box, sign, v.x, v.y, iv.x, iv.y, P.dt, rng, Eb);
                if (nslots_of<P1>(P) > 1 && alive) f0 = first_pick<P1>(P, rng, mode, box);
            }
        }
        scatter_out<P1>(P, H, sbox, vtab, rng, f0, alive, mode, sign, box, p.x, p.y, p.pid, step, &lcount, myseg);
        done += alive ? 1 : 0;
    }
    seg_close(P, segout, myseg, &lcount);
    eb_lds_flush(P, Eg, nb4);
    box_lds_flush(P, myseg);
    // one atomic per workgroup for the particle-step counter
    block_add_u64(P.stats, done);
}

// Emission sampling: first index i with cdf[i] >= u (the reference's
// lower_bound on the cumulative), started from a guide table.
#define BERT_GUIDE 1024
__device__ inline int guide_key(double v, double scale) {
    const double k = v * scale;
    return k >= (double)(BERT_GUIDE - 1) ? BERT_GUIDE - 1 : (int)k;
}

// inclusive prefix sum of v over the 256 threads of a workgroup (wave scans
// by shuffles, then the four wave totals through LDS); total in *tot
__device__ inline double block_scan_256(double v, double* wtot, double& tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        const double u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) wtot[w] = v;
    __syncthreads();
    double off = 0.;
    for (int k = 0; k < w; k++) off += wtot[k];
    tot = wtot[0] + wtot[1] + wtot[2] + wtot[3];
    __syncthreads();  // wtot is reused by the next call
    return v + off;
}

// inclusive max-scan over the 256 threads of a workgroup (wave scans by
// shuffles, the four wave maxima through LDS)
__device__ inline int block_max_scan_256(int v, int* wmax) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v = max(v, u);
    }
    if (lane == 63) wmax[w] = v;
    __syncthreads();
    for (int k = 0; k < w; k++) v = max(v, wmax[k]);
    return v;
}

#define BERT_WCDF_PER 8
// Adiabatic diffuse walls (beyondRTA_MC2d.cpp:281-435), phase 1: one
// workgroup per (box, side) with a nonzero Eborder computes the border
// temperature from the J(T) spline and the emission cumulative of
// Isothermal_wall_distribution(Tborder, Teq, -n) (sampling.cpp:189-269).
__global__ __launch_bounds__(256) void bert_wallcdf_kernel(BertParams P, uint32_t step) {
    __shared__ double wtot[4];
    __shared__ int kcarry[4];
    // register path: the weight of every mode, then its guide key
    __shared__ double spm[256 * BERT_WCDF_PER];
    __shared__ int skey[256 * BERT_WCDF_PER];
    __shared__ int grun[BERT_GUIDE];  // mode starting the run of keys at each guide entry, else 0
    const int b = blockIdx.x;  // side
    const int n = P.n;
    const bool regs = n <= 256 * BERT_WCDF_PER;
    // The kernel is a chain of dependent loads (side -> box -> spline ->
    // modes), so the mode velocities and frequencies, which depend on
    // nothing, are requested first, coalesced: mode k * 256 + thread.
    double2 mv[BERT_WCDF_PER];
    double mom[BERT_WCDF_PER];
    if (regs) {
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            const int i = k * 256 + threadIdx.x;
            mv[k] = i < n ? P.vel[i] : make_double2(0., 0.);
            mom[k] = i < n ? P.omega[i] : 0.;
        }
    }
    const int box = P.side_box[b];
    int32_t* Eb = P.Eb + (size_t)(step & 1) * P.nsd;
    const int E = Eb[b];
    const int si = P.side_spl[b];
    if (E == 0 || si < 0) {
        if (threadIdx.x == 0) {
            P.wcount[b] = 0;
            Eb[b] = 0;
        }
        return;
    }
    __syncthreads();
    if (threadIdx.x == 0) Eb[b] = 0;
    const DevSide sd = P.sides[b];
    const double Teq = P.box_Teq[box];
    const double* beq = P.beq + (size_t)P.box_beq[box] * n;
    double mbq[BERT_WCDF_PER];
    if (regs) {
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            const int i = k * 256 + threadIdx.x;
            mbq[k] = i < n ? beq[i] : 0.;
        }
    }
    const double L = hypot(sd.p2x - sd.p1x, sd.p2y - sd.p1y);
    const double J = P.Eeff * (double)E / P.dt / (L * P.thickness);
    const double* S = P.spl + (size_t)si * 6 * P.nseg;  // a b c d x0 xe
    // lower bound of J among the ascending knots: the count of knots below J
    int lo = 0;
    for (int k0 = 0; k0 < P.nseg; k0 += blockDim.x) {
        const int k = k0 + threadIdx.x;
        lo += __syncthreads_count(k < P.nseg && S[5 * P.nseg + k] < J);
    }
    if (lo >= P.nseg) lo = P.nseg - 1;
    const double dx = J - S[4 * P.nseg + lo];
    const double Tb = S[lo] + dx * (S[P.nseg + lo] + dx * (S[2 * P.nseg + lo] + S[3 * P.nseg + lo] * dx));
    const double nx = -sd.nx, ny = -sd.ny;  // emission normal points into the box
    // n(omega, Tb) - n(omega, Teq): the first from exp(omega pf / Tb), the
    // second from the box's row of the equilibrium table
    constexpr double pf = 1e12 * hbar / kB;
    const double sTb = pf / Tb;
    double* cdf = P.wcdf + (size_t)b * n;
    int32_t* G = P.wguide + (size_t)b * BERT_GUIDE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (regs) {
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            const int i = k * 256 + threadIdx.x;
            const double vn = nx * mv[k].x + ny * mv[k].y;
            double pm = 0.;
            if (vn > 0. && !almost_equal(0., mom[k]))
                pm = vn * mom[k] * fabs(1. / (exp(sTb * mom[k]) - 1.) - mbq[k]);
            if (i < n) spm[i] = pm;
        }
        for (int j = threadIdx.x; j < BERT_GUIDE; j += blockDim.x) grun[j] = 0;
        __syncthreads();
        // the scan: each thread sums up to BERT_WCDF_PER consecutive modes
        // from LDS, one block scan adds the preceding threads' totals
        const int per = (n + 255) >> 8;
        const int i0 = threadIdx.x * per;
        double c[BERT_WCDF_PER];
        double run = 0.;
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            if (k < per && i0 + k < n) run += spm[i0 + k];
            c[k] = run;
        }
        double tot;
        const double excl = block_scan_256(run, wtot, tot) - run;
        const double inv = tot > 0. ? 1.0 / tot : 0.;
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            if (k < per && i0 + k < n) {
                c[k] = (c[k] + excl) * inv;
                spm[i0 + k] = c[k];
                skey[i0 + k] = guide_key(c[k], (double)BERT_GUIDE);
            }
        }
        if (threadIdx.x == 0) {
            P.wcount[b] = tot > 0. ? (uint32_t)(fabs((double)E) + 1.0e-6) : 0u;
            P.wsign[b] = (Tb - Teq) >= 0 ? 1 : -1;
        }
        if (!(tot > 0.)) return;  // uniform over the workgroup
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) cdf[i] = spm[i];  // coalesced
        // G[j] = first mode whose key reaches j (the keys ascend with the
        // cumulative), the last mode when none does. Mode i owns the keys
        // (key[i-1], key[i]]: it marks the first of them, the last mode marks
        // the key above its own, and a max-scan over the entries carries each
        // mark up to the next one (the marking modes ascend with the entries).
#pragma unroll
        for (int k = 0; k < BERT_WCDF_PER; k++) {
            const int i = i0 + k;
            if (k < per && i < n) {
                const int kp = i > 0 ? skey[i - 1] : -1;
                if (skey[i] > kp) grun[kp + 1] = i;
            }
        }
        if (threadIdx.x == 0 && skey[n - 1] + 1 < BERT_GUIDE) grun[skey[n - 1] + 1] = n - 1;
        __syncthreads();
        constexpr int GPT = BERT_GUIDE / 256;  // entries per thread
        int gv[GPT];
        int gm = 0;
#pragma unroll
        for (int q = 0; q < GPT; q++) {
            gm = max(gm, grun[threadIdx.x * GPT + q]);
            gv[q] = gm;
        }
        const int gpre = block_max_scan_256(gm, kcarry);
        // the entries before this thread's: its predecessor's scan value
        // (for a wave's first lane, the maximum of the earlier waves)
        int gprev = __shfl_up(gpre, 1, 64);
        if (lane == 0) {
            gprev = 0;
            for (int k = 0; k < w; k++) gprev = max(gprev, kcarry[k]);
        }
        static_assert(GPT == 4, "one 16-byte store of guide entries per thread");
        *(int4*)(G + threadIdx.x * GPT) = make_int4(max(gv[0], gprev), max(gv[1], gprev), max(gv[2], gprev), max(gv[3], gprev));
        return;
    }
    // larger tables: the cumulative in chunks of 256 consecutive modes, one mode per thread
    // (coalesced), carrying the running total from chunk to chunk
    double carry = 0.;
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int i = c0 + threadIdx.x;
        double pm = 0.;
        if (i < n) {
            const double2 v = P.vel[i];
            const double om = P.omega[i];
            const double vn = nx * v.x + ny * v.y;
            if (vn > 0. && !almost_equal(0., om)) pm = vn * om * fabs(1. / (exp(sTb * om) - 1.) - beq[i]);
        }
        double ctot;
        const double incl = block_scan_256(pm, wtot, ctot);
        if (i < n) cdf[i] = carry + incl;
        carry += ctot;
    }
    const double tot = carry;
    if (threadIdx.x == 0) {
        P.wcount[b] = tot > 0. ? (uint32_t)(fabs((double)E) + 1.0e-6) : 0u;
        P.wsign[b] = (Tb - Teq) >= 0 ? 1 : -1;
    }
    if (!(tot > 0.)) return;
    // guide table: G[j] = first index whose key reaches j, so the search for
    // u starts at G[key(u)] <= answer (key is monotone); each thread fills
    // the keys between its predecessor's and its own
    const double inv = 1.0 / tot;
    int kprev_chunk = -1;  // key of the last mode of the previous chunk
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int i = c0 + threadIdx.x;
        // own cdf values were written by this thread: no barrier needed;
        // normalised to a last value of 1
        double cn = 0.;
        if (i < n) {
            cn = cdf[i] * inv;
            cdf[i] = cn;
        }
        const int k = i < n ? guide_key(cn, (double)BERT_GUIDE) : -1;
        int kp = __shfl_up(k, 1, 64);
        if (lane == 63) kcarry[w] = k;
        __syncthreads();
        if (lane == 0) kp = w > 0 ? kcarry[w - 1] : kprev_chunk;
        const int klast = kcarry[3];
        __syncthreads();
        if (i < n)
            for (int j = kp + 1; j <= k; j++) G[j] = i;
        kprev_chunk = klast;
    }
    // keys above the last cumulative value (rounding) point at the last mode
    // (cdf[n-1] was written before the barriers above)
    const int klast = guide_key(cdf[n - 1], (double)BERT_GUIDE);
    for (int j = klast + 1 + threadIdx.x; j < BERT_GUIDE; j += blockDim.x) G[j] = n - 1;
}

// Phase 2: the |Eborder| particles of every side form one index space split
// in equal slices over the workgroups; each is placed uniformly on its side,
// advected U*dt, scattered and histogrammed. 512-thread workgroups (two per
// CU) keep four waves per SIMD resident with one output segment per
// workgroup: 256-thread ones left two (and more segments slow the step
// kernel's segment walk).
#define BERT_EMIT_MAXB 4096
#define BERT_EMIT_THREADS 512
// VL: as in the step kernel, one 1024-thread workgroup per CU with the mode
// velocities and their inverses in LDS.
template <bool GL, bool PER, bool P1, bool POLY, bool VL = false>
__global__ __launch_bounds__(VL ? 1024 : BERT_EMIT_THREADS) void bert_emit_kernel(BertParams P, uint32_t step, int cur,
                                                                                  int hist) {
    __shared__ uint32_t tsum[16];
    __shared__ uint32_t lcount;
    uint32_t* pre = (uint32_t*)(bert_dyn + P.lds_pre);
    const int nb4 = P.nsd;
    const int myseg = P.seg_emit0 + blockIdx.x;
    uint32_t* segout = P.seg + (size_t)(cur ^ 1) * P.nsegs;
    if (threadIdx.x == 0) lcount = 0;
    const DevBox* boxes;
    const DevContact* cts;
    geo_stage<GL>(P, boxes, cts);
    const double2* vtab = P.vel;
    const double2* ivtab = P.ivel;
    if constexpr (VL) {
        // visible after seg_prefix's barrier below
        double2* sv = (double2*)(bert_dyn + P.lds_vel);
        for (int q = threadIdx.x; q < P.n; q += blockDim.x) {
            sv[q] = P.vel[q];
            sv[P.n + q] = P.ivel[q];
        }
        vtab = sv;
        ivtab = sv + P.n;
    }
    int32_t* Eg = P.Eb + (size_t)((step + 1) & 1) * nb4;
    int32_t* Ebn = eb_lds_begin(P, Eg, nb4);
    double* sbox = box_lds_begin(P);
    const uint32_t total = seg_prefix(P.wcount, nb4, 0xffffffffu, pre, tsum);  // ends with a barrier
    const uint32_t lo = (uint32_t)((uint64_t)total * blockIdx.x / gridDim.x);
    const uint32_t hi = (uint32_t)((uint64_t)total * (blockIdx.x + 1) / gridDim.x);
    unsigned long long* H = hist ? P.H + (size_t)cur * P.nbox * P.n : nullptr;
    const uint64_t skey = stream_key(P.seed, step, 1);
    unsigned long long done = 0;
    int bw = lo < hi ? seg_search(pre, nb4, lo) : 0;
    for (uint32_t base = lo; base < hi; base += blockDim.x) {
        const uint32_t k = base + threadIdx.x;
        const bool valid = k < hi;
        const uint32_t wfirst = base + (threadIdx.x & ~63u);
        while (bw < nb4 - 1 && wfirst >= pre[bw]) bw++;
        Pcg32 rng;
        Pick0 f0{0u, 0u, make_uint2(0u, 0u)};
        int mode = 0, bx = 0, sign = 1;
        double x = 0, y = 0;
        bool alive = false;
        uint64_t pid = 0;
        if (valid) {
            // side b holds indices [pre[b-1], pre[b])
            int b = bw;
            while (b < nb4 - 1 && k >= pre[b]) b++;
            const uint32_t j = k - (b > 0 ? pre[b - 1] : 0u);
            bx = POLY ? P.side_box[b] : b >> 2;
            const int side = b & 3;
            sign = P.wsign[b];
            pid = mix64(P.seed ^ ((uint64_t)step << 40) ^ ((uint64_t)b << 20) ^ (uint64_t)j ^ 0xad1aULL);
            seed_stream(rng, skey, pid);
            // normalised cumulative: no load of the side's total before the guide
            const double* cdf = P.wcdf + (size_t)b * P.n;
            const double u = rng.uniform();
            int l = P.wguide[(size_t)b * BERT_GUIDE + guide_key(u, (double)BERT_GUIDE)];
            // first l at or after the guide with cdf[l] >= u (the last mode when
            // none): four entries per round of loads instead of one
            const int nl = P.n - 1;
            for (;;) {
                const int l1 = min(l + 1, nl), l2 = min(l + 2, nl), l3 = min(l + 3, nl);
                const double c0 = cdf[l], c1 = cdf[l1], c2 = cdf[l2], c3 = cdf[l3];
                if (l >= nl || c0 >= u) break;
                if (l1 >= nl || c1 >= u) { l = l1; break; }
                if (l2 >= nl || c2 >= u) { l = l2; break; }
                if (l3 >= nl || c3 >= u) { l = l3; break; }
                l += 4;
            }
            mode = l;
            double p1x, p1y, ndx, ndy;
            if constexpr (POLY) {
                const DevSide sd = P.sides[b];
                p1x = sd.p1x; p1y = sd.p1y; ndx = sd.p2x - sd.p1x; ndy = sd.p2y - sd.p1y;
            } else {
                const DevBox B = boxes[bx];
                if (side == 0) { p1x = B.xmin; p1y = B.ymin; ndx = 0; ndy = B.ymax - B.ymin; }
                else if (side == 1) { p1x = B.xmin; p1y = B.ymin; ndx = B.xmax - B.xmin; ndy = 0; }
                else if (side == 2) { p1x = B.xmin; p1y = B.ymax; ndx = B.xmax - B.xmin; ndy = 0; }
                else { p1x = B.xmax; p1y = B.ymax; ndx = 0; ndy = B.ymin - B.ymax; }
            }
            double t = rng.uniform();
            x = p1x + t * ndx;
            y = p1y + t * ndy;
            const double2 v = vtab[mode];
            const double2 iv = ivtab[mode];
            double rdt = P.dt * rng.uniform();
            if (nslots_of<P1>(P) == 1) f0 = first_pick<P1>(P, rng, mode, 0);
            alive = advect<PER, POLY>(P, boxes, cts, x, y, bx, sign, v.x, v.y, iv.x, iv.y, rdt, rng, Ebn);
            if (nslots_of<P1>(P) > 1 && alive) f0 = first_pick<P1>(P, rng, mode, bx);
        }
        scatter_out<P1>(P, H, sbox, vtab, rng, f0, alive, mode, sign, bx, x, y, pid, step, &lcount, myseg);
        done += alive ? 1 : 0;
    }
    seg_close(P, segout, myseg, &lcount);
    eb_lds_flush(P, Eg, nb4);
    box_lds_flush(P, myseg);
    block_add_u64(P.stats + 1, done);
}

// Fixed-rate sources, one or more parts each with a fixed alias table:
// reservoir injection (Isothermal_walls, beyondRTA_MC2d.cpp:446-585) on a
// contact segment p1 + t nd with a fixed sign, and gradient sources
// (source_term_gradient, 597-676) at uniform points of a box (area: p1 the
// lower corner, nd the box size) with the sign in the alias entry. At most
// BERT_SRC_BLOCKS workgroups run the parts, each a contiguous range of
// them into its own buffer segment, so the segment count does not grow
// with the number of sources.
struct BertSource {
    double p1x, p1y, ndx, ndy;
    double expected;  // particles per step
    int32_t box, sign;
    int32_t tab;      // alias table offset (in entries)
    int32_t src;      // source index (keys its random streams)
    int32_t part, nparts;  // this part's share of the source's particles
    int32_t area, pad;
};
#define BERT_SRC_BLOCKS 256
template <bool PER, bool POLY>
__global__ __launch_bounds__(256) void bert_source_kernel(BertParams P, const BertSource* src, int nparts_all,
                                                          const uint2* stab, uint32_t step, int cur, int hist) {
    __shared__ uint32_t lcount;
    const int myseg = P.seg_src0 + blockIdx.x;
    uint32_t* segout = P.seg + (size_t)(cur ^ 1) * P.nsegs;
    if (threadIdx.x == 0) lcount = 0;
    double* sbox = box_lds_begin(P);
    __syncthreads();
    int32_t* Eb = P.Eb + (size_t)(step & 1) * P.nsd;
    unsigned long long* H = hist ? P.H + (size_t)cur * P.nbox * P.n : nullptr;
    unsigned long long done = 0;
    const int j0 = (int)((int64_t)nparts_all * blockIdx.x / gridDim.x);
    const int j1 = (int)((int64_t)nparts_all * (blockIdx.x + 1) / gridDim.x);
    for (int j = j0; j < j1; j++) {
        const BertSource S = src[j];
        // every part of a source draws the same count (Ntogenerate,
        // sampling.cpp:230-245) and takes its slice of the particle indices
        Pcg32 cr;
        seed_stream(cr, stream_key(P.seed, step, 2), 0x5ce0000000000000ULL + (uint64_t)S.src);
        const uint64_t skey = stream_key(P.seed, step, 3);
        const double fl = floor(S.expected);
        const int ng = (int)fl + (cr.uniform() <= S.expected - fl ? 1 : 0);
        const int klo = (int)((int64_t)ng * S.part / S.nparts), khi = (int)((int64_t)ng * (S.part + 1) / S.nparts);
        const int nloop = (khi - klo + 63) & ~63;
        for (int q = threadIdx.x; q - (int)(threadIdx.x & 63) < nloop; q += blockDim.x) {
            const int k = klo + q;
            bool valid = k < khi;
            Pcg32 rng;
            Pick0 f0{0u, 0u, make_uint2(0u, 0u)};
            int mode = 0, bx = S.box, sign = S.sign;
            double x = 0, y = 0;
            bool alive = false;
            uint64_t pid = mix64(P.seed ^ ((uint64_t)step << 40) ^ ((uint64_t)S.src << 20) ^ (uint64_t)k ^ 0x15a0ULL);
            if (valid) {
                seed_stream(rng, skey, pid);
                double t = rng.uniform();
                if (S.area) {
                    // get_random_point (geometry_2d.hpp:332-348): x, then y
                    // in the bounding box until covered by the box
                    x = S.p1x + t * S.ndx;
                    y = S.p1y + rng.uniform() * S.ndy;
                    if constexpr (POLY) {
                        const int2 sr = P.srange[S.box];
                        int tries = 0;
                        while (!poly_inside(P.sides, sr, x, y)) {
                            if (++tries > 4096) {  // a sliver in its bounding box: reported
                                atomicOr(P.flags + 3, 1u);
                                x = P.center[S.box].x;
                                y = P.center[S.box].y;
                                break;
                            }
                            x = S.p1x + rng.uniform() * S.ndx;
                            y = S.p1y + rng.uniform() * S.ndy;
                        }
                    }
                } else {
                    x = S.p1x + t * S.ndx;
                    y = S.p1y + t * S.ndy;
                    const DevBox B = P.boxes[S.box];
                    const bool in = POLY ? poly_inside(P.sides, P.srange[S.box], x, y) : box_inside(B, x, y);
                    if (!in) {
                        const double ccx = POLY ? P.center[S.box].x : 0.5 * (B.xmin + B.xmax);
                        const double ccy = POLY ? P.center[S.box].y : 0.5 * (B.ymin + B.ymax);
                        double dx = ccx - x, dy = ccy - y;
                        double dn = sqrt(dx * dx + dy * dy);
                        x += 1.0e-6 * dx / dn;
                        y += 1.0e-6 * dy / dn;
                    }
                }
                bool ngf;
                mode = (int)alias_pick(stab + S.tab, (uint32_t)P.n, rng, ngf);
                if (S.area) sign = ngf ? -1 : 1;
                const double2 v = P.vel[mode];
                const double2 iv = P.ivel[mode];
                double rdt = P.dt * rng.uniform();
                if (nslots_of<false>(P) == 1) f0 = first_pick<false>(P, rng, mode, 0);
                alive = advect<PER, POLY>(P, P.boxes, P.cts, x, y, bx, sign, v.x, v.y, iv.x, iv.y, rdt, rng, Eb);
                if (nslots_of<false>(P) > 1 && alive) f0 = first_pick<false>(P, rng, mode, bx);
            }
            scatter_out<false>(P, H, sbox, P.vel, rng, f0, alive, mode, sign, bx, x, y, pid, step, &lcount, myseg);
            done += alive ? 1 : 0;
        }
    }
    seg_close(P, segout, myseg, &lcount);
    box_lds_flush(P, myseg);
    block_add_u64(P.stats + 1, done);
}

// The i-th of n workgroups, dispatched round-robin over 8 XCDs, takes item
// start(x) + i / 8 on XCD x = i % 8: each XCD gets a contiguous range.
__device__ inline int xcd_major_index(int i, int n) {
    const int x = i & 7, j = i >> 3;
    int start = 0;
    for (int y = 0; y < x; y++) start += (n - y + 7) >> 3;
    return start + j;
}

// calculate_prop (beyondRTA_MC2d.cpp:1535-1602) from the step's box sums
// (from_boxes) or its histogram, the post-cancellation particle count, and
// buffer recycling for the next step (clear_old: the previous step wrote its
// histogram, which the step kernel has now consumed).
__global__ __launch_bounds__(256) void bert_props_kernel(BertParams P, int row, int cur, int cancel_now,
                                                         int clear_tickets, int ebn, int from_boxes,
                                                         int clear_old) {
    // Workgroups go to the 8 XCDs round-robin; consecutive boxes are given
    // to workgroups on the same XCD, so the cache lines of pbox rows (a
    // 128-B line holds the sums of five neighbouring boxes) are fetched
    // once into that XCD's L2 instead of once per box.
    const int box = xcd_major_index((int)blockIdx.x, (int)gridDim.x);
    const unsigned long long* H = P.H + (size_t)cur * P.nbox * P.n + (size_t)box * P.n;
    unsigned long long* Hold = P.H + (size_t)(cur ^ 1) * P.nbox * P.n + (size_t)box * P.n;
    __shared__ double red[3][256];
    __shared__ long long redc[2][256];
    double s = 0, fx = 0, fy = 0;
    long long cnt = 0, acnt = 0;
    if (!from_boxes || cancel_now || clear_old || clear_tickets) {
        for (int k = threadIdx.x; k < P.n; k += blockDim.x) {
            if (!from_boxes || cancel_now) {
                const int h = h_net(H[k]);
                if (h != 0) {
                    const double2 v = P.vel[k];
                    s += h;
                    fx += h * v.x;
                    fy += h * v.y;
                    cnt += h;
                    acnt += h < 0 ? -h : h;
                }
            }
            if (clear_old) Hold[k] = 0;
            if (clear_tickets) P.tickets[(size_t)box * P.n + k] = 0;
        }
    }
    if (from_boxes) {
        // fixed summation order over the writing workgroups
        s = fx = fy = 0.;
        for (int g = threadIdx.x; g < P.nsegs; g += blockDim.x) {
            const double* q = P.pbox + ((size_t)g * P.nbox + box) * 3;
            s += q[0];
            fx += q[1];
            fy += q[2];
        }
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = fx;
    red[2][threadIdx.x] = fy;
    redc[0][threadIdx.x] = cnt;
    redc[1][threadIdx.x] = acnt;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) {
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + d];
            redc[0][threadIdx.x] += redc[0][threadIdx.x + d];
            redc[1][threadIdx.x] += redc[1][threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const DevBox B = P.boxes[box];
        double vol = P.box_area[box] * P.thickness;
        bool skip = B.kind != 0;
        double* o = P.props + ((size_t)row * P.nbox + box) * 3;
        o[0] = skip ? 0. : P.Eeff / vol * 1.0e27 * red[0][0];
        o[1] = skip ? 0. : P.Eeff / vol * 1.0e30 * red[1][0];
        o[2] = skip ? 0. : P.Eeff / vol * 1.0e30 * red[2][0];
        // particles left after this step: sum |h| over regular boxes when
        // the step cancels, else every particle written this step
        if (cancel_now && !skip)
            atomicAdd((unsigned long long*)(P.npart + row), (unsigned long long)redc[1][0]);
        const int2 sr = P.srange[box];
        const int32_t* E = P.Eb + (size_t)ebn * P.nsd + sr.x;
        long long pend = 0;
        for (int k = 0; k < sr.y; k++) pend += E[k];
        if (pend) atomicAdd((unsigned long long*)(P.pending + row), (unsigned long long)pend);
    }
    if (box == 0) {
        // live particles written this step; recycle the input counters
        __syncthreads();
        uint32_t* segin = P.seg + (size_t)cur * P.nsegs;
        const uint32_t* segw = P.seg + (size_t)(cur ^ 1) * P.nsegs;
        long long c = 0;
        for (int k = threadIdx.x; k < P.nsegs; k += blockDim.x) {
            c += min(segw[k], P.segcap);
            segin[k] = 0;
        }
        redc[0][threadIdx.x] = c;
        __syncthreads();
        for (int d = 128; d > 0; d >>= 1) {
            if (threadIdx.x < d) redc[0][threadIdx.x] += redc[0][threadIdx.x + d];
            __syncthreads();
        }
        if (threadIdx.x == 0 && !cancel_now) P.npart[row] = (int64_t)redc[0][0];
    }
}

// initial particles: histogram into H[cur] (properties row 0)
__global__ __launch_bounds__(256) void bert_hist_kernel(BertParams P, int cur) {
    __shared__ uint32_t pre[BERT_MAXSEG];
    __shared__ uint32_t tsum[16];
    const uint32_t nin = seg_prefix(P.seg + (size_t)cur * P.nsegs, P.nsegs, P.segcap, pre, tsum);
    unsigned long long* H = P.H + (size_t)cur * P.nbox * P.n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nin;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const int sg = seg_search(pre, P.nsegs, (uint32_t)i);
        const BPart p = P.pin[(uint64_t)sg * P.segcap + ((uint32_t)i - (sg > 0 ? pre[sg - 1] : 0u))];
        int box = (int)(p.boxsign >> 1);
        int sign = (p.boxsign & 1u) ? -1 : 1;
        h_add(H + (size_t)box * P.n + p.mode, sign);
    }
}

// initial particles from outTref distributions (beyondRTA_MC2d.cpp:211-269):
// one workgroup per box with its count and alias table; particle pid goes to
// segment pid % nsegs
struct BertInit {
    double xmin, xmax, ymin, ymax;
    uint64_t first, count;
    int32_t box, sign, tab, pad;
};
__global__ __launch_bounds__(256) void bert_init_kernel(BertParams P, const BertInit* ini,
                                                        const uint2* itab) {
    const BertInit I = ini[blockIdx.x];
    for (uint64_t k = (uint64_t)blockIdx.y * blockDim.x + threadIdx.x; k < I.count;
         k += (uint64_t)gridDim.y * blockDim.x) {
        uint64_t pid = I.first + k;
        Pcg32 rng;
        seed_stream(rng, stream_key(P.seed, 0xffffffu, 4), pid);
        double x = I.xmin + rng.uniform() * (I.xmax - I.xmin);
        double y = I.ymin + rng.uniform() * (I.ymax - I.ymin);
        if (P.poly) {
            // get_random_point: the bounding box until covered by the box
            const int2 sr = P.srange[I.box];
            int tries = 0;
            while (!poly_inside(P.sides, sr, x, y)) {
                if (++tries > 4096) {  // a sliver in its bounding box: reported
                    atomicOr(P.flags + 3, 1u);
                    x = P.center[I.box].x;
                    y = P.center[I.box].y;
                    break;
                }
                x = I.xmin + rng.uniform() * (I.xmax - I.xmin);
                y = I.ymin + rng.uniform() * (I.ymax - I.ymin);
            }
        }
        bool ng;
        int mode = (int)alias_pick(itab + I.tab, (uint32_t)P.n, rng, ng);
        // tables of given distributions carry each mode's sign in the alias word
        const int sg = ng ? -I.sign : I.sign;
        BPart q{x, y, (uint32_t)mode, ((uint32_t)I.box << 1) | (sg < 0 ? 1u : 0u), pid};
        P.pin[(pid % P.nsegs) * P.segcap + pid / P.nsegs] = q;
    }
}

// particle kernels by [geometry in LDS][periodic boxes][P1]; polygon
// geometries by [periodic boxes][P1] (geometry in global memory)
using StepKernel = void (*)(BertParams, uint32_t, int, int, int);
using EmitKernel = void (*)(BertParams, uint32_t, int, int);
static const StepKernel step_kernel_tab[2][2][2] = {
    {{bert_step_kernel<false, false, false, false>, bert_step_kernel<false, false, true, false>},
     {bert_step_kernel<false, true, false, false>, bert_step_kernel<false, true, true, false>}},
    {{bert_step_kernel<true, false, false, false>, bert_step_kernel<true, false, true, false>},
     {bert_step_kernel<true, true, false, false>, bert_step_kernel<true, true, true, false>}}};
static const EmitKernel emit_kernel_tab[2][2][2] = {
    {{bert_emit_kernel<false, false, false, false>, bert_emit_kernel<false, false, true, false>},
     {bert_emit_kernel<false, true, false, false>, bert_emit_kernel<false, true, true, false>}},
    {{bert_emit_kernel<true, false, false, false>, bert_emit_kernel<true, false, true, false>},
     {bert_emit_kernel<true, true, false, false>, bert_emit_kernel<true, true, true, false>}}};
static const StepKernel step_kernel_poly_tab[2][2] = {
    {bert_step_kernel<false, false, false, true>, bert_step_kernel<false, false, true, true>},
    {bert_step_kernel<false, true, false, true>, bert_step_kernel<false, true, true, true>}};
// the VL step kernels (no periodic boxes): [GL][P1] for rectangles, [P1] for polygons
static const StepKernel step_kernel_vl_tab[2][2] = {
    {bert_step_kernel<false, false, false, false, true>, bert_step_kernel<false, false, true, false, true>},
    {bert_step_kernel<true, false, false, false, true>, bert_step_kernel<true, false, true, false, true>}};
static const StepKernel step_kernel_poly_vl_tab[2] = {bert_step_kernel<false, false, false, true, true>,
                                                       bert_step_kernel<false, false, true, true, true>};
static const EmitKernel emit_kernel_vl_tab[2][2] = {
    {bert_emit_kernel<false, false, false, false, true>, bert_emit_kernel<false, false, true, false, true>},
    {bert_emit_kernel<true, false, false, false, true>, bert_emit_kernel<true, false, true, false, true>}};
static const EmitKernel emit_kernel_poly_vl_tab[2] = {bert_emit_kernel<false, false, false, true, true>,
                                                       bert_emit_kernel<false, false, true, true, true>};
static const EmitKernel emit_kernel_poly_tab[2][2] = {
    {bert_emit_kernel<false, false, false, true>, bert_emit_kernel<false, false, true, true>},
    {bert_emit_kernel<false, true, false, true>, bert_emit_kernel<false, true, true, true>}};

struct Bert {
    Ctx* ctx = nullptr;
    const Modes* m = nullptr;
    const Geom* g = nullptr;
    const Prop* p = nullptr;
    bte_bert_cfg cfg{};
    double Eeff = 0;
    int max_steps = 0;
    uint64_t n0 = 0, cap = 0;
    int nseg = 99;
    DevBuf<double2> d_vel, d_ivel;
    DevBuf<BPart> d_parts;
    DevBuf<uint32_t> d_seg, d_tickets, d_flags;
    DevBuf<int32_t> d_Eb, d_side_spl, d_rta_slot, d_side_box;
    DevBuf<unsigned long long> d_H;
    DevBuf<double> d_props, d_spl, d_Teq, d_pbox, d_box_area, d_beq;
    DevBuf<int32_t> d_box_beq;
    DevBuf<double2> d_center;
    DevBuf<int64_t> d_npart, d_pending;
    DevBuf<double> d_wcdf;
    DevBuf<int32_t> d_wguide;
    DevBuf<uint32_t> d_wcount;
    DevBuf<int32_t> d_wsign;
    DevBuf<unsigned long long> d_stats;
    DevBuf<uint2> d_rta_alias, d_src_tab, d_init_tab;
    DevBuf<BertSource> d_src;
    DevBuf<BertInit> d_init;
    // propagator slots: one table per regular box temperature
    DevBuf<uint2> d_palias;
    DevBuf<int32_t> d_pslot;
    const uint2* palias = nullptr;
    int nslots = 1;
    int nsrc = 0, src_blocks = 0, ninit = 0;  // source parts, workgroups running them
    bool geo_lds = false;
    bool vl = false;        // step kernels with the mode velocities in LDS (one 1024-thread workgroup per CU)
    size_t lds_vl_bytes = 0;
    bool periodic = false;  // some box is periodic: launch the PER kernels
    bool p1 = false;        // expected to launch the P1 kernels (sizes the step grid)
    size_t lds_bytes = 0;
    int step_blocks = 0, emit_blocks = 0, nsegs = 0;
    std::vector<int32_t> hH;
    std::vector<unsigned long long> hH64;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> kev;  // begin/end event per step-kernel launch
    ~Bert() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto e : kev) (void)hipEventDestroy(e);
    }
};

static double calc_Eeff_host(const Modes& m, double Th, double Tc, double Tr, double vol, uint64_t np) {
    long double dh = 0, dc = 0;
    for (int i = 0; i < m.n; i++) {
        if (almost_equal(0., m.omega[i])) continue;
        dh += 1e12 * hbar * m.omega[i] * std::fabs(h_bose_einstein(m.omega[i], Th) - h_bose_einstein(m.omega[i], Tr));
        dc += 1e12 * hbar * m.omega[i] * std::fabs(h_bose_einstein(m.omega[i], Tc) - h_bose_einstein(m.omega[i], Tr));
    }
    return (double)(vol * (dc + dh) / m.nq / np);
}

// The pair of tabulated temperatures for a box temperature Tx that is not
// tabulated. rule 0 is Scattering's choice (beyondRTA_MC2d.cpp:1139-1164):
// T1 = the first key not below Tx, T2 = the key after it, allowed when T2
// exists and Tx >= the lowest key (the line through T1, T2 is evaluated at
// Tx < T1). rule 1 brackets Tx between adjacent keys.
static void prop_pair(const std::vector<double>& keys, double Tx, int rule, int& t1, int& t2) {
    const int i = (int)(std::lower_bound(keys.begin(), keys.end(), Tx) - keys.begin());
    const int nk = (int)keys.size();
    if (rule == 0) {
        if (i < nk && i + 1 < nk && Tx >= keys.front()) {
            t1 = i;
            t2 = i + 1;
            return;
        }
    } else if (i > 0 && i < nk) {
        t1 = i - 1;
        t2 = i;
        return;
    }
    throw Error(BTE_EINVAL, "Extrapolation is not allowed: box temperature " + std::to_string(Tx) +
                                " must lie between the min and max temperatures used in Propagator calculation");
}

// Alias tables (sign bits in the alias word) for an untabulated temperature.
// rule 0: lirp(H(T1), H(T2), Tx) for every column, the interpolated signed
// cumulative sampled as propagator_H::get samples it (pick_pmf), sign =
// (sc > 0 ? + : -). rule 1: the column of (1 - f) P(T1) + f P(T2), sampled by
// its absolute values with its signs. Interpolating signed cumulatives is
// unstable where an entry's sign differs between the two tables (rounding
// noise on negligible entries does that): the cumulative then dips and the
// next entry takes a large, possibly negative, probability.
static void interp_table(const Prop& p, int t1, int t2, double Tx, int rule, std::vector<uint2>& tab) {
    const int n = p.ncol;
    tab.resize((size_t)n * n);
    const double T1 = p.T[t1], T2 = p.T[t2];
    const double f = (Tx - T1) / (T2 - T1);
    auto work = [&](int j0, int j1) {
        std::vector<double> sc1(n), sc2(n), sc(n), pmf(n);
        std::vector<uint32_t> th(n), al(n);
        auto signed_cumm = [n](const double* col, double* out) {
            double tot = 0;
            for (int k = 0; k < n; k++) tot += std::fabs(col[k]);
            double acc = 0;
            for (int k = 0; k < n; k++) {
                acc += std::fabs(col[k]);
                out[k] = col[k] < 0. ? -acc / tot : acc / tot;
            }
        };
        for (int j = j0; j < j1; j++) {
            const double* c1 = p.P[t1].data() + (size_t)j * n;
            const double* c2 = p.P[t2].data() + (size_t)j * n;
            if (rule == 1) {
                // sc holds the interpolated column itself: only its signs are used
                for (int k = 0; k < n; k++) {
                    sc[k] = c1[k] + f * (c2[k] - c1[k]);
                    pmf[k] = std::fabs(sc[k]);
                }
            } else {
                signed_cumm(c1, sc1.data());
                signed_cumm(c2, sc2.data());
                for (int k = 0; k < n; k++) sc[k] = sc1[k] + f * (sc2[k] - sc1[k]);
                pick_pmf(sc.data(), n, pmf.data());
            }
            if (!build_alias(pmf.data(), n, th.data(), al.data()))
                throw Error(BTE_EINVAL, "interpolated propagator column " + std::to_string(j) + " is zero");
            for (int k = 0; k < n; k++) {
                const uint32_t a = al[k];
                const uint32_t word = a | (sc[k] > 0. ? 0u : 0x80000000u) | (sc[a] > 0. ? 0u : 0x40000000u);
                tab[(size_t)j * n + k] = make_uint2(th[k], word);
            }
        }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> ths;
    std::vector<std::exception_ptr> errs(nt);
    for (int w = 0; w < nt; w++)
        ths.emplace_back([&, w] {
            try {
                work((int)((int64_t)n * w / nt), (int)((int64_t)n * (w + 1) / nt));
            } catch (...) {
                errs[w] = std::current_exception();
            }
        });
    for (auto& t : ths) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

static BertParams make_params(Bert* b) {
    BertParams P{};
    P.vel = b->d_vel.p;
    P.ivel = b->d_ivel.p;
    P.omega = b->m->d_omega.p;
    P.w = b->m->d_w.p;
    P.boxes = b->g->d_boxes.p;
    P.cts = b->g->d_contacts.p;
    P.cseg = b->g->d_cseg.p;
    P.ptrans = b->g->d_ptrans.p;
    P.sides = b->g->d_sides.p;
    P.srange = b->g->d_srange.p;
    P.side_box = b->d_side_box.p;
    P.center = b->d_center.p;
    P.box_area = b->d_box_area.p;
    P.nsd = (int32_t)b->g->sides.size();
    P.poly = b->g->poly ? 1 : 0;
    P.palias = b->palias;
    P.box_pslot = b->d_pslot.p;
    P.nslots = b->nslots;
    P.rta_alias = b->d_rta_alias.p;
    P.box_rta_slot = b->d_rta_slot.p;
    P.box_Teq = b->d_Teq.p;
    P.beq = b->d_beq.p;
    P.box_beq = b->d_box_beq.p;
    P.spl = b->d_spl.p;
    P.side_spl = b->d_side_spl.p;
    P.seg = b->d_seg.p;
    P.H = b->d_H.p;
    P.tickets = b->d_tickets.p;
    P.Eb = b->d_Eb.p;
    P.props = b->d_props.p;
    P.npart = b->d_npart.p;
    P.pending = b->d_pending.p;
    P.wcdf = b->d_wcdf.p;
    P.wguide = b->d_wguide.p;
    P.wcount = b->d_wcount.p;
    P.wsign = b->d_wsign.p;
    P.flags = b->d_flags.p;
    P.stats = b->d_stats.p;
    P.cap = b->cap;
    P.nsegs = b->nsegs;
    P.segcap = (uint32_t)(b->cap / b->nsegs);
    P.seg_emit0 = b->step_blocks;
    P.seg_src0 = b->step_blocks + b->emit_blocks;
    P.ncts = (int)b->g->dcontacts.size();
    P.seed = b->cfg.seed;
    P.dt = b->cfg.dt;
    P.Eeff = b->Eeff;
    P.thickness = b->cfg.thickness;
    P.n = b->m->n;
    P.ngamma = b->p ? b->p->ngamma : 0;
    P.ncol = b->p ? b->p->ncol : 0;
    P.nbox = (int)b->g->boxes.size();
    P.nseg = b->nseg;
    P.rta = b->cfg.rta;
    // a particle's sign-flip branching stops with an error after this many
    // children (a table with negative probability >= 1/2 never ends)
    P.max_children = 1 << 14;
    P.pbox = b->d_pbox.p;
    P.lds_props = (int)b->g->boxes.size() <= BERT_LDS_BOXES ? 1 : 0;
    // dynamic LDS layout
    const int nb4 = P.nsd;
    auto up16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
    size_t off = 0;
    P.lds_pre = (int32_t)off;
    off = up16(off + sizeof(uint32_t) * (size_t)std::max(P.nsegs, nb4));
    P.eb_lds = nb4 <= BERT_EB_LDS ? 1 : 0;
    P.lds_eb = (int32_t)off;
    if (P.eb_lds) off = up16(off + sizeof(int32_t) * (size_t)nb4);
    P.lds_sbox = (int32_t)off;
    if (P.lds_props) off = up16(off + sizeof(double) * 3 * (size_t)P.nbox);
    const size_t geo = sizeof(DevBox) * (size_t)P.nbox + sizeof(DevContact) * (size_t)P.ncts;
    b->geo_lds = geo <= BERT_GEO_LDS_BYTES && !b->g->poly;  // polygon kernels read the geometry globally
    P.lds_boxes = (int32_t)off;
    P.lds_cts = (int32_t)up16(off + sizeof(DevBox) * (size_t)P.nbox);
    P.cts_lds = 1;
    if (b->geo_lds) {
        off = up16(P.lds_cts + sizeof(DevContact) * (size_t)P.ncts);
        // the step workgroups share the CU's LDS: five 256-thread ones, or
        // one VL workgroup with the velocity tables after the layout. When
        // the layout does not fit that share, the contacts (read only for
        // sides that several boxes share and for periodic translations) stay
        // in global memory
        const long long per_wg = b->vl ? (long long)b->ctx->lds_per_block - 1024 - 32 * (long long)P.n
                                       : (long long)(b->ctx->lds_per_block / 5);
        if ((b->vl || b->p1) && !b->periodic && (long long)off + 256 > per_wg && (long long)P.lds_cts + 256 <= per_wg) {
            P.cts_lds = 0;
            off = (size_t)P.lds_cts;
        }
    }
    b->lds_bytes = off;
    P.lds_vel = (int32_t)up16(off);
    b->lds_vl_bytes = up16(off) + 32 * (size_t)P.n;
    // the particle kernels' dynamic LDS and their static arrays (under 1 KB)
    if (off + 1024 > b->ctx->lds_per_block)
        throw Error(BTE_EINVAL, "the particle kernels need " + std::to_string(off + 1024) + " bytes of LDS for " +
                                    std::to_string(P.nbox) + " boxes, " + std::to_string(P.nsd) + " sides and " +
                                    std::to_string(P.nsegs) + " buffer segments; the device has " +
                                    std::to_string(b->ctx->lds_per_block) + " per workgroup");
    return P;
}

}  // namespace bte

using namespace bte;

extern "C" {

int bte_prop_create_multi(bte_ctx* ctx, int ncol, int ntemp, const double* Ps, const double* Ts, int ngamma,
                          bte_prop** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !Ps || !Ts || !out || ncol <= 0 || ntemp <= 0 || ngamma < 0 || ncol >= (1 << 30))
        throw Error(BTE_EINVAL, "bad arguments to bte_prop_create_multi");
    HIPCHECK(hipSetDevice(c->device));
    auto* p = new Prop();
    std::unique_ptr<Prop> guard(p);
    p->ctx = c;
    p->ncol = ncol;
    p->ngamma = ngamma;
    // std::map<double, hist_type> semantics (beyondRTA_MC2d.cpp:1352-1378):
    // ascending temperatures, a repeated temperature keeps the last matrix
    const size_t nn = (size_t)ncol * ncol;
    std::map<double, int> last;
    for (int t = 0; t < ntemp; t++) {
        if (!std::isfinite(Ts[t])) throw Error(BTE_EINVAL, "propagator temperature is not finite");
        last[Ts[t]] = t;
    }
    if (last.size() > 1 && last.begin()->first <= 0.)
        throw Error(BTE_EINVAL, "propagator temperatures must be positive");
    for (auto& [T, t] : last) {
        p->T.push_back(T);
        p->P.emplace_back(Ps + (size_t)t * nn, Ps + (size_t)(t + 1) * nn);
    }
    // propagator_H per column (collision_operator.cpp:854-870) as an alias
    // table over |P_kj| with the entry signs in the alias word
    std::vector<uint2> tab(p->T.size() * nn);
    std::vector<double> pm(ncol);
    std::vector<uint32_t> th(ncol), al(ncol);
    for (size_t t = 0; t < p->T.size(); t++) {
        for (int j = 0; j < ncol; j++) {
            const double* col = p->P[t].data() + (size_t)j * ncol;
            for (int k = 0; k < ncol; k++) pm[k] = std::fabs(col[k]);
            if (!build_alias(pm.data(), ncol, th.data(), al.data()))
                throw Error(BTE_EINVAL, "propagator column " + std::to_string(j) + " is zero");
            for (int k = 0; k < ncol; k++) {
                uint32_t a = al[k];
                uint32_t word = a | (col[k] < 0. ? 0x80000000u : 0u) | (col[a] < 0. ? 0x40000000u : 0u);
                tab[t * nn + (size_t)j * ncol + k] = make_uint2(th[k], word);
            }
        }
    }
    p->d_alias.upload(tab, c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
    *out = reinterpret_cast<bte_prop*>(guard.release());
    BTE_GUARD_END
}

int bte_prop_create(bte_ctx* ctx, int ncol, const double* P, int ngamma, double T, bte_prop** out) {
    return bte_prop_create_multi(ctx, ncol, 1, P, &T, ngamma, out);
}

int bte_prop_pick_pmf(int n, const double* signed_cumm, double* pmf) {
    BTE_GUARD_BEGIN
    if (n <= 0 || !signed_cumm || !pmf) throw Error(BTE_EINVAL, "bad arguments to bte_prop_pick_pmf");
    pick_pmf(signed_cumm, n, pmf);
    BTE_GUARD_END
}

int bte_geom_gradient_T(const bte_geom* gg, double* grad) {
    BTE_GUARD_BEGIN
    auto* g = reinterpret_cast<const Geom*>(gg);
    if (!g || !grad) throw Error(BTE_EINVAL, "null argument");
    auto gr = gradient_T(*g);
    for (size_t i = 0; i < gr.size(); i++) {
        grad[2 * i] = gr[i].x;
        grad[2 * i + 1] = gr[i].y;
    }
    BTE_GUARD_END
}

int bte_prop_destroy(bte_prop* pp) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<Prop*>(pp);
    BTE_GUARD_END
}

int bte_bert_create(bte_ctx* ctx, const bte_modes* mm, const bte_geom* gg, const bte_prop* pp,
                    const bte_bert_cfg* cfg, bte_bert** out) {
    BTE_GUARD_BEGIN
    auto* c = reinterpret_cast<Ctx*>(ctx);
    auto* m = reinterpret_cast<const Modes*>(mm);
    auto* g = reinterpret_cast<const Geom*>(gg);
    auto* p = reinterpret_cast<const Prop*>(pp);
    if (!c || !m || !g || !cfg || !out) throw Error(BTE_EINVAL, "null argument");
    if (!cfg->rta && !p) throw Error(BTE_EINVAL, "a propagator is required unless rta is set");
    if (p && p->ngamma + p->ncol != m->n) throw Error(BTE_EINVAL, "propagator size does not match the mode table");
    if (!(cfg->dt > 0) || !(cfg->thickness > 0)) throw Error(BTE_EINVAL, "bad beRTA configuration");
    if (g->sides.size() > BERT_EMIT_MAXB)
        throw Error(BTE_EINVAL, "more than " + std::to_string(BERT_EMIT_MAXB) + " box sides");
    for (size_t i = 0; i < g->boxes.size(); i++)
        if (g->dboxes[i].kind == 2 && g->dboxes[i].tgt < 0)
            throw Error(BTE_EGEOM, "periodic box " + std::to_string(i) + " has no translation");
    HIPCHECK(hipSetDevice(c->device));
    auto* b = new Bert();
    std::unique_ptr<Bert> guard(b);
    b->ctx = c;
    b->m = m;
    b->g = g;
    b->p = cfg->rta ? nullptr : p;
    b->cfg = *cfg;
    if (b->cfg.pcancel_each <= 0) b->cfg.pcancel_each = 4;
    hipStream_t s = c->stream;
    const int n = m->n, nbox = (int)g->boxes.size();
    double vuc = m->V * cfg->thickness / m->lattz;
    // steps: while (time <= maxtime) (beyondRTA_MC2d.cpp:2138)
    int nst = 0;
    {
        double t = 0;
        while (t <= cfg->maxtime) { nst++; t += cfg->dt; }
    }
    b->max_steps = cfg->max_steps > 0 ? std::min(cfg->max_steps, nst) : nst;
    // Eeff (beyondRTA_MC2d.cpp:2026-2070)
    if (cfg->Ed < 0.) {
        double maxT = 0, minT = 1.416784e+32, meanT = 0;
        int nb_ = 0;
        for (auto& bx : g->boxes) {
            if (bx.kind == 2) continue;
            meanT += bx.Teq;
            nb_++;
            minT = std::min(minT, bx.Teq);
            maxT = std::max(maxT, bx.Teq);
        }
        if (cfg->Tmax != 0) { maxT = cfg->Tmax; minT = cfg->Tmin; }
        meanT /= nb_;
        if (cfg->Efactor == 0) throw Error(BTE_EINVAL, "Eeff needs Ed or particles");
        double e = 1e300;
        for (auto& bx : g->boxes) {
            if (bx.kind != 0) continue;
            e = std::min(e, calc_Eeff_host(*m, maxT, minT, meanT, bx.area() * cfg->thickness / vuc, cfg->Efactor));
        }
        b->Eeff = e;
    } else {
        b->Eeff = cfg->Ed;
    }
    if (!(b->Eeff > 0)) throw Error(BTE_EINVAL, "particle energy Eeff must be positive");
    // initial particles (outTref per box with Tinit != Teq), counts drawn on
    // the host like Ntogenerate (sampling.cpp:313-325)
    Pcg32 hr;
    hr.seed(cfg->seed, 0xb0a7u);
    std::vector<BertInit> inits;
    std::vector<uint2> itab;
    std::vector<double> pmf(n);
    std::vector<uint32_t> th(n), al(n);
    uint64_t total = 0;
    // given distributions first (init_particles, beyondRTA_MC2d.cpp:142-193;
    // ref_distribution, sampling.cpp:70-114): pmf |fd| over the modes after
    // the Gamma offset, the sign of fd in the alias word, Ntogenerate(Vbox /
    // Vucell, Eeff) = vol sum|fd| / (Eeff nq) particles
    if (cfg->ninit > 0) {
        const int ngam = b->p ? b->p->ngamma : 0;
        if (!cfg->init_boxes || !cfg->init_dist) throw Error(BTE_EINVAL, "init_distribution without data");
        if (cfg->init_len + ngam != n)
            throw Error(BTE_EINVAL, "init_distribution holds " + std::to_string(cfg->init_len) + " entries, the modes "
                                        "after the Gamma offset are " + std::to_string(n - ngam));
        std::vector<int8_t> neg(n);
        for (int q = 0; q < cfg->ninit; q++) {
            const int i = cfg->init_boxes[q];
            if (i < 0 || i >= nbox) throw Error(BTE_EINVAL, "init_distribution box out of range");
            const RectBox& bx = g->boxes[i];
            if (bx.kind != 0) continue;  // reservoirs and periodic boxes are skipped
            const double* fd = cfg->init_dist + (size_t)q * cfg->init_len;
            double cs = 0;
            for (int k = 0; k < n; k++) {
                pmf[k] = k < ngam ? 0. : std::fabs(fd[k - ngam]);
                neg[k] = k >= ngam && std::signbit(fd[k - ngam]) ? 1 : 0;
                if (!std::isfinite(pmf[k])) throw Error(BTE_EINVAL, "init_distribution entries must be finite");
                cs += pmf[k];
            }
            const double gen = (bx.area() * cfg->thickness / vuc) * cs / (b->Eeff * m->nq);
            uint64_t ng = (uint64_t)gen;
            if (hr.uniform() <= gen - std::floor(gen)) ng++;
            if (!(cs > 0.)) continue;
            build_alias(pmf.data(), n, th.data(), al.data());
            BertInit I{bx.xmin, bx.xmax, bx.ymin, bx.ymax, total, ng, i, 1, (int32_t)itab.size(), 0};
            for (int k = 0; k < n; k++) {
                const uint32_t a = al[k];
                itab.push_back(make_uint2(th[k], a | (neg[k] ? 0x80000000u : 0u) | (neg[a] ? 0x40000000u : 0u)));
            }
            inits.push_back(I);
            total += ng;
        }
    }
    for (int i = 0; i < nbox; i++) {
        const RectBox& bx = g->boxes[i];
        if (bx.kind != 0 || almost_equal(bx.Teq, bx.Tinit)) continue;
        if (bx.Tinit <= 0 || bx.Teq <= 0) throw Error(BTE_EINVAL, "Invalid temperature value");
        double acc = 0;
        for (int k = 0; k < n; k++) {
            pmf[k] = almost_equal(m->omega[k], 0.) ? 0. :
                m->omega[k] * std::fabs(h_bose_einstein(m->omega[k], bx.Tinit) - h_bose_einstein(m->omega[k], bx.Teq));
            acc += pmf[k];
        }
        double cumulsum = 1.0e12 * hbar * acc;
        double gen = (bx.area() * cfg->thickness / vuc) * cumulsum / (b->Eeff * m->nq);
        uint64_t ng = (uint64_t)gen;
        if (hr.uniform() <= gen - std::floor(gen)) ng++;
        build_alias(pmf.data(), n, th.data(), al.data());
        BertInit I{bx.xmin, bx.xmax, bx.ymin, bx.ymax, total, ng, i, (bx.Tinit - bx.Teq) >= 0 ? 1 : -1,
                   (int32_t)itab.size(), 0};
        for (int k = 0; k < n; k++) itab.push_back(make_uint2(th[k], al[k]));
        inits.push_back(I);
        total += ng;
    }
    b->n0 = total;
    b->ninit = (int)inits.size();
    // one buffer segment per writing workgroup: step, emission, sources
    // reservoir sources (Isothermal_walls): expected particles per step,
    // split over workgroups of at most ~1024 particles each
    std::vector<BertSource> srcs;
    std::vector<uint2> stab;
    double src_total = 0;
    for (int sid = 0; sid < nbox; sid++) {
        const RectBox& sb = g->boxes[sid];
        if (sb.kind != 1) continue;
        for (auto& hc : g->contacts[sid]) {
            const RectBox& cb = g->boxes[hc.box];
            if (cb.kind != 0 || almost_equal(hc.length(), 0.) || almost_equal(sb.Teq, cb.Teq)) continue;
            double cum = wall_pmf(*m, sb.Teq, cb.Teq, hc.npx, hc.npy, pmf.data());
            double spf = hc.length() * cfg->thickness / vuc;
            build_alias(pmf.data(), n, th.data(), al.data());
            const double expected = cfg->dt * spf * cum / (b->Eeff * m->nq);
            const int nparts = (int)std::min(256.0, std::max(1.0, std::ceil(expected / 1024.0)));
            const int src = (int)(stab.size() / n);
            for (int part = 0; part < nparts; part++)
                srcs.push_back(BertSource{hc.p1x, hc.p1y, hc.p2x - hc.p1x, hc.p2y - hc.p1y, expected, hc.box,
                                          (sb.Teq - cb.Teq) >= 0 ? 1 : -1, (int32_t)stab.size(), src, part, nparts,
                                          0, 0});
            for (int k = 0; k < n; k++) stab.push_back(make_uint2(th[k], al[k]));
            src_total += expected;
        }
    }
    // gradient sources (source_term_gradient, beyondRTA_MC2d.cpp:597-676):
    // per regular box with a gradient that is not almost zero, the
    // Nabla_T_distribution (sampling.cpp:118-164) at the box Teq, pmf
    // |(-g.v) kernel(omega, Teq)| with the sign of -g.v in the alias word,
    // Ntogenerate(dt, vol, Eeff) (sampling.cpp:166-183) particles per step
    {
        std::vector<double2> grad;
        if (cfg->fixed_gradient) {
            grad.assign(nbox, make_double2(0., 0.));
            for (int i = 0; i < nbox; i++)
                if (g->boxes[i].kind == 0) grad[i] = make_double2(cfg->gradient[0], cfg->gradient[1]);
        } else {
            grad = gradient_T(*g);
        }
        std::vector<int8_t> neg(n);
        for (int i = 0; i < nbox; i++) {
            const RectBox& bx = g->boxes[i];
            if (bx.kind != 0 || almost_equal(std::hypot(grad[i].x, grad[i].y), 0.)) continue;
            double acc = 0;
            for (int k = 0; k < n; k++) {
                pmf[k] = 0.;
                neg[k] = 0;
                if (m->omega[k] == 0.) continue;
                const double vgt = -(grad[i].x * m->vx[k] + grad[i].y * m->vy[k]);
                pmf[k] = std::fabs(vgt * h_be_kernel(m->omega[k], bx.Teq));
                neg[k] = vgt >= 0. ? 0 : 1;
                acc += pmf[k];
            }
            if (!(acc > 0.)) continue;  // no mode carries this gradient
            const double cumulsum = kB * acc;
            build_alias(pmf.data(), n, th.data(), al.data());
            const double expected = cfg->dt * (bx.area() * cfg->thickness / vuc) * cumulsum / (b->Eeff * m->nq);
            const int nparts = (int)std::min(256.0, std::max(1.0, std::ceil(expected / 1024.0)));
            const int src = (int)(stab.size() / n);
            for (int part = 0; part < nparts; part++)
                srcs.push_back(BertSource{bx.xmin, bx.ymin, bx.xmax - bx.xmin, bx.ymax - bx.ymin, expected, i, 1,
                                          (int32_t)stab.size(), src, part, nparts, 1, 0});
            for (int k = 0; k < n; k++) {
                const uint32_t a = al[k];
                stab.push_back(make_uint2(th[k], a | (neg[k] ? 0x80000000u : 0u) | (neg[a] ? 0x40000000u : 0u)));
            }
            src_total += expected;
        }
    }
    b->nsrc = (int)srcs.size();
    // step workgroups per CU: as many as its registers keep resident (3 with
    // the periodic translation compiled in, 5 for the P1 kernels without it,
    // 4 otherwise), for latency hiding. P1: no RTA and one propagator table,
    // i.e. a single tabulated temperature or one Teq over the regular boxes
    for (auto& bx : g->dboxes) b->periodic = b->periodic || bx.kind == 2;
    bool one_table = true;
    if (b->p && !(b->p->T.size() == 1 && b->p->T[0] <= 0.)) {
        std::set<double> temps;
        for (auto& bx : g->boxes)
            if (bx.kind == 0) temps.insert(bx.Teq);
        one_table = temps.size() <= 1;
    }
    b->p1 = !cfg->rta && b->p != nullptr && one_table;
    // step, emission and source workgroups together stay within
    // BERT_MAXSEG buffer segments on any device
    b->src_blocks = std::min(b->nsrc, BERT_SRC_BLOCKS);
    // the VL step and emission kernels (one 1024-thread workgroup per CU,
    // mode velocities in LDS) whenever no box is periodic and the tables fit
    // next to the particle kernels' layout; else 256-thread step and
    // 512-thread emission workgroups
    if (cfg->step_shape < 0 || cfg->step_shape > 1) throw Error(BTE_EINVAL, "step_shape must be 0 or 1");
    b->vl = !b->periodic && cfg->step_shape == 0;
    if (b->vl) {
        b->step_blocks = std::min(c->cu_count, BERT_MAXSEG - 512 - BERT_SRC_BLOCKS);
        b->emit_blocks = std::min(c->cu_count, 512);
        b->nsegs = b->step_blocks + b->emit_blocks + b->src_blocks;
        (void)make_params(b);  // the layout for these segments
        b->vl = b->lds_vl_bytes + 1024 <= c->lds_per_block;
    }
    if (!b->vl) {
        b->step_blocks = std::min((b->periodic ? 3 : (b->p1 ? 5 : 4)) * c->cu_count, BERT_MAXSEG - 512 - BERT_SRC_BLOCKS);
        b->emit_blocks = std::min(2 * c->cu_count, 512);
    }
    b->nsegs = b->step_blocks + b->emit_blocks + b->src_blocks;
    if (b->nsegs > BERT_MAXSEG) throw Error(BTE_EINVAL, "too many buffer segments");
    // generous by default (32 B per particle and buffer): reservoir-driven
    // runs grow to a steady population, and every segment keeps >= 4096 slots
    // the busiest source workgroup's expected particles per step, twice over,
    // fit one segment
    double src_wg_max = 0;
    for (int w = 0; w < b->src_blocks; w++) {
        double e = 0;
        const int j0 = (int)((int64_t)b->nsrc * w / b->src_blocks), j1 = (int)((int64_t)b->nsrc * (w + 1) / b->src_blocks);
        for (int j = j0; j < j1; j++) e += srcs[j].expected / srcs[j].nparts;
        src_wg_max = std::max(src_wg_max, e);
    }
    // (every step workgroup's segment takes twice its share of the initial
    // particles; with sources, its share of a steady population of 0.7 x 2^25)
    const uint64_t seg_step = std::max<uint64_t>(2 * total, src_total > 0 ? (uint64_t)(0.7 * (1ull << 25)) : 0ull) /
                                  (uint64_t)b->step_blocks + 4096;
    b->cap = cfg->capacity ? cfg->capacity
                           : std::max<uint64_t>({1ull << 22, 3 * total, src_total > 0 ? (1ull << 25) : 0ull,
                                                 4096ull * (uint64_t)b->nsegs, seg_step * (uint64_t)b->nsegs,
                                                 (uint64_t)(2.0 * src_wg_max + 4096.0) * (uint64_t)b->nsegs});
    b->cap = (b->cap + b->nsegs - 1) / b->nsegs * b->nsegs;
    if (b->cap >= 0xffffffffULL) throw Error(BTE_EINVAL, "capacity above 2^32 particles");
    if ((total + b->nsegs - 1) / b->nsegs > b->cap / b->nsegs)
        throw Error(BTE_ECAPACITY, "initial particles exceed the capacity");
    // device buffers
    std::vector<double2> vel(n);
    for (int k = 0; k < n; k++) vel[k] = make_double2(m->vx[k], m->vy[k]);
    b->d_vel.upload(vel, s);
    std::vector<double2> ivel(n);
    for (int k = 0; k < n; k++)
        ivel[k] = make_double2(m->vx[k] != 0. ? 1.0 / m->vx[k] : 0., m->vy[k] != 0. ? 1.0 / m->vy[k] : 0.);
    b->d_ivel.upload(ivel, s);
    b->d_parts.alloc(2 * b->cap);
    b->d_seg.alloc(2 * (size_t)b->nsegs);
    b->d_flags.alloc(4);
    b->d_stats.alloc(2);
    b->d_H.alloc((size_t)2 * nbox * n);
    b->d_tickets.alloc((size_t)nbox * n);
    if (nbox <= BERT_LDS_BOXES) b->d_pbox.alloc((size_t)b->nsegs * nbox * 3);
    b->d_Eb.alloc((size_t)2 * g->sides.size());
    b->d_props.alloc((size_t)(b->max_steps + 1) * nbox * 3);
    b->d_npart.alloc((size_t)(b->max_steps + 1));
    b->d_pending.alloc((size_t)(b->max_steps + 1));
    const size_t nsd = g->sides.size();
    b->d_wcdf.alloc(nsd * n);
    b->d_wguide.alloc(nsd * BERT_GUIDE);
    b->d_wcount.alloc(nsd);
    b->d_wsign.alloc(nsd);
    std::vector<double> Teq(nbox), area(nbox);
    std::vector<double2> centre(nbox);
    std::vector<int32_t> side_box(nsd);
    for (int i = 0; i < nbox; i++) {
        Teq[i] = g->boxes[i].Teq;
        area[i] = g->boxes[i].area();
        centre[i] = make_double2(g->boxes[i].cx(), g->boxes[i].cy());
        for (int k = 0; k < g->srange[i].y; k++) side_box[g->srange[i].x + k] = i;
    }
    b->d_Teq.upload(Teq, s);
    {
        // equilibrium occupations per distinct box temperature
        std::map<double, int> rows;
        std::vector<double> beq;
        std::vector<int32_t> box_beq(nbox, 0);
        for (int i = 0; i < nbox; i++) {
            const double T = g->boxes[i].Teq;
            if (!(T > 0.)) continue;  // no walls emit there
            if (!rows.count(T)) {
                rows[T] = (int)rows.size();
                for (int k = 0; k < n; k++) beq.push_back(m->omega[k] > 0. ? h_bose_einstein(m->omega[k], T) : 0.);
            }
            box_beq[i] = rows[T];
        }
        if (beq.empty()) beq.assign((size_t)n, 0.);
        b->d_beq.upload(beq, s);
        b->d_box_beq.upload(box_beq, s);
    }
    b->d_box_area.upload(area, s);
    b->d_center.upload(centre, s);
    b->d_side_box.upload(side_box, s);
    if (!inits.empty()) {
        b->d_init.upload(inits, s);
        b->d_init_tab.upload(itab, s);
    }
    // RTA samplers per distinct box temperature (ScatteringRTA)
    std::vector<int32_t> slot(nbox, 0);
    if (cfg->rta) {
        std::map<double, int> slots;
        std::vector<uint2> rt;
        for (int i = 0; i < nbox; i++) {
            double T = g->boxes[i].Teq;
            if (!slots.count(T)) {
                int id = (int)slots.size();
                slots[T] = id;
                for (int k = 0; k < n; k++) pmf[k] = m->omega[k] <= 0. ? 0. : h_be_kernel(m->omega[k], T) * m->w[k];
                build_alias(pmf.data(), n, th.data(), al.data());
                for (int k = 0; k < n; k++) rt.push_back(make_uint2(th[k], al[k]));
            }
            slot[i] = slots[T];
        }
        b->d_rta_alias.upload(rt, s);
    } else {
        b->d_rta_alias.alloc(1);
    }
    b->d_rta_slot.upload(slot, s);
    // propagator per regular box temperature (Scattering,
    // beyondRTA_MC2d.cpp:1131-1166): the tabulated one, or the lirp
    // (collision_operator.cpp:918-929) of the pair prop_pair picks; every
    // regular box is checked here, before the first step
    std::vector<int32_t> pslot(nbox, 0);
    b->nslots = 1;
    if (b->p) {
        const Prop* pp = b->p;
        const size_t nn = (size_t)pp->ncol * pp->ncol;
        if (pp->T.size() == 1 && pp->T[0] <= 0.) {
            b->palias = pp->d_alias.p;
        } else {
            struct SlotSrc { int exact, t1, t2; double Tx; };
            std::map<double, int> slot_of;
            std::vector<SlotSrc> ss;
            for (int i = 0; i < nbox; i++) {
                const double T = g->boxes[i].Teq;
                if (g->boxes[i].kind != 0) continue;
                if (!slot_of.count(T)) {
                    SlotSrc q{-1, 0, 0, T};
                    for (size_t t = 0; t < pp->T.size(); t++)
                        if (pp->T[t] == T) q.exact = (int)t;
                    if (q.exact < 0) prop_pair(pp->T, T, cfg->prop_interp, q.t1, q.t2);
                    slot_of[T] = (int)ss.size();
                    ss.push_back(q);
                }
                pslot[i] = slot_of[T];
            }
            if (ss.empty()) ss.push_back(SlotSrc{0, 0, 0, pp->T[0]});
            b->nslots = (int)ss.size();
            if (b->nslots == 1 && ss[0].exact >= 0) {
                b->palias = pp->d_alias.p + (size_t)ss[0].exact * nn;
            } else {
                b->d_palias.alloc(ss.size() * nn);
                std::vector<uint2> tab;
                for (size_t q = 0; q < ss.size(); q++) {
                    if (ss[q].exact >= 0) {
                        HIPCHECK(hipMemcpyAsync(b->d_palias.p + q * nn, pp->d_alias.p + (size_t)ss[q].exact * nn,
                                                nn * sizeof(uint2), hipMemcpyDeviceToDevice, s));
                    } else {
                        interp_table(*pp, ss[q].t1, ss[q].t2, ss[q].Tx, cfg->prop_interp, tab);
                        HIPCHECK(hipMemcpyAsync(b->d_palias.p + q * nn, tab.data(), nn * sizeof(uint2),
                                                hipMemcpyHostToDevice, s));
                        HIPCHECK(hipStreamSynchronize(s));  // tab is reused
                    }
                }
                b->palias = b->d_palias.p;
            }
        }
    }
    b->d_pslot.upload(pslot, s);
    // adiabatic wall splines J(T), T = 10..1000 K (beyondRTA_MC2d.cpp:294-350)
    std::vector<double> Ts(100);
    for (int i = 0; i < 100; i++) Ts[i] = 10.0 + 10.0 * i;
    std::map<std::pair<double, int>, int> spl_id;
    std::vector<double> spl;
    std::vector<int32_t> side_spl(g->sides.size(), -1);
    for (int i = 0; i < nbox; i++) {
        const RectBox& bx = g->boxes[i];
        if (bx.kind != 0) continue;
        for (int sd = g->srange[i].x; sd < g->srange[i].x + g->srange[i].y; sd++) {
            // one J(T) table per box temperature and side normal
            auto key = std::make_pair(bx.Teq, (int)g->sides[sd].nid);
            if (!spl_id.count(key)) {
                const double nx = g->sides[sd].nx, ny = g->sides[sd].ny;
                std::vector<double> J(100);
                for (int t = 0; t < 100; t++) {
                    double cum = wall_pmf(*m, Ts[t], bx.Teq, -nx, -ny, pmf.data());
                    int sg = (Ts[t] - bx.Teq) >= 0 ? 1 : -1;
                    J[t] = sg * cum / (m->nq * vuc);
                }
                Spline sp;
                sp.build(J, Ts);
                int id = (int)spl_id.size();
                spl_id[key] = id;
                for (auto* v : {&sp.a, &sp.b, &sp.c, &sp.d, &sp.x0, &sp.xe}) spl.insert(spl.end(), v->begin(), v->end());
            }
            side_spl[sd] = spl_id[key];
        }
    }
    b->d_spl.upload(spl.empty() ? std::vector<double>(1, 0.) : spl, s);
    b->d_side_spl.upload(side_spl, s);
    if (b->nsrc) {
        b->d_src.upload(srcs, s);
        b->d_src_tab.upload(stab, s);
    }
    b->hH.resize((size_t)nbox * n);
    b->hH64.resize((size_t)nbox * n);
    b->kev.resize((size_t)2 * b->max_steps);
    for (auto& e : b->kev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipEventCreate(&b->ev0));
    HIPCHECK(hipEventCreate(&b->ev1));
    (void)make_params(b);  // refuses a layout that exceeds the LDS before any run
    HIPCHECK(hipStreamSynchronize(s));
    *out = reinterpret_cast<bte_bert*>(guard.release());
    BTE_GUARD_END
}

int bte_bert_execute(bte_bert* bb, double* props, int64_t* npart, int64_t* pending, bte_frame_cb cb,
                     void* user, bte_bert_report* rep) {
    BTE_GUARD_BEGIN
    auto* b = reinterpret_cast<Bert*>(bb);
    if (!b || !props || !npart) throw Error(BTE_EINVAL, "null argument");
    Ctx* c = b->ctx;
    HIPCHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int n = b->m->n, nbox = (int)b->g->boxes.size();
    BertParams P = make_params(b);
    BPart* buf[2] = {b->d_parts.p, b->d_parts.p + b->cap};
    HIPCHECK(hipMemsetAsync(b->d_H.p, 0, (size_t)2 * nbox * n * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(b->d_tickets.p, 0, (size_t)nbox * n * sizeof(uint32_t), s));
    HIPCHECK(hipMemsetAsync(b->d_Eb.p, 0, (size_t)2 * b->g->sides.size() * sizeof(int32_t), s));
    HIPCHECK(hipMemsetAsync(b->d_flags.p, 0, 4 * sizeof(uint32_t), s));
    HIPCHECK(hipMemsetAsync(b->d_stats.p, 0, 2 * sizeof(unsigned long long), s));
    HIPCHECK(hipMemsetAsync(b->d_npart.p, 0, (size_t)(b->max_steps + 1) * sizeof(int64_t), s));
    HIPCHECK(hipMemsetAsync(b->d_pending.p, 0, (size_t)(b->max_steps + 1) * sizeof(int64_t), s));
    std::vector<uint32_t> c0(2 * (size_t)b->nsegs, 0u);
    for (int k = 0; k < b->nsegs; k++)
        c0[k] = (uint32_t)(b->n0 / b->nsegs + ((uint64_t)k < b->n0 % b->nsegs ? 1 : 0));
    HIPCHECK(hipMemcpyAsync(b->d_seg.p, c0.data(), c0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    P.pin = buf[0];
    P.pout = buf[1];
    if (b->ninit) {
        hipLaunchKernelGGL(bert_init_kernel, dim3(b->ninit, 1024), dim3(256), 0, s, P, b->d_init.p,
                           b->d_init_tab.p);
        HIPCHECK(hipGetLastError());
    }
    // row 0: properties of the initial particles
    hipLaunchKernelGGL(bert_hist_kernel, dim3(c->cu_count * 4), dim3(256), 0, s, P, 0);
    HIPCHECK(hipGetLastError());
    // props of H[0] into row 0; the kernel clears H[1] and the input segment
    // counters; restore them afterwards because the initial particles are
    // the first step's input
    hipLaunchKernelGGL(bert_props_kernel, dim3(nbox), dim3(256), 0, s, P, 0, 0, 0, 0, 0, 0, 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(b->d_seg.p, c0.data(), c0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(b->d_H.p, 0, (size_t)nbox * n * sizeof(unsigned long long), s));
    HIPCHECK(hipMemcpyAsync(b->d_npart.p, &b->n0, sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    const int step_blocks = b->step_blocks;
    const int emit_blocks = b->emit_blocks;
    // one propagator table and no RTA scattering: the P1 particle kernels
    const int p1 = (!b->cfg.rta && P.nslots == 1) ? 1 : 0;
    const int pc = b->cfg.pcancel_each;
    StepKernel step_k = b->g->poly ? step_kernel_poly_tab[b->periodic][p1] : step_kernel_tab[b->geo_lds][b->periodic][p1];
    EmitKernel emit_k = b->g->poly ? emit_kernel_poly_tab[b->periodic][p1] : emit_kernel_tab[b->geo_lds][b->periodic][p1];
    if (b->vl) {
        step_k = b->g->poly ? step_kernel_poly_vl_tab[p1] : step_kernel_vl_tab[b->geo_lds][p1];
        emit_k = b->g->poly ? emit_kernel_poly_vl_tab[p1] : emit_kernel_vl_tab[b->geo_lds][p1];
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(step_k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)b->lds_vl_bytes));
        HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(emit_k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)b->lds_vl_bytes));
    }
    auto t0 = std::chrono::steady_clock::now();
    HIPCHECK(hipEventRecord(b->ev0, s));
    int istep = 0;
    int prev_hist = 0;
    for (; istep < b->max_steps; istep++) {
        const int cur = istep & 1;
        P.pin = buf[cur];
        P.pout = buf[cur ^ 1];
        const int cancel_prev = istep > 0 && ((istep - 1) % pc == 0);
        const int cancel_now = (istep % pc == 0);
        // the (box, mode) histogram is built only where a cancellation or a
        // frame reads it; properties come from the box sums otherwise
        const int hist = cancel_now || cb != nullptr || !P.lds_props;
        // one launch in pc carries an event pair, rotating through the
        // cancellation cycle (steps pc*q + (q mod pc)): an event record
        // between two kernels leaves the device idle for ~5.6 us
        const bool timed = (istep % pc) == ((istep / pc) % pc);
        if (timed) HIPCHECK(hipEventRecord(b->kev[2 * istep], s));
        hipLaunchKernelGGL(step_k, dim3(step_blocks), dim3(b->vl ? 1024 : 256), b->vl ? b->lds_vl_bytes : b->lds_bytes,
                           s, P, (uint32_t)istep, cur, cancel_prev, hist);
        if (timed) HIPCHECK(hipEventRecord(b->kev[2 * istep + 1], s));
        if (b->nsrc)
            hipLaunchKernelGGL(b->g->poly ? (b->periodic ? bert_source_kernel<true, true> : bert_source_kernel<false, true>)
                                          : (b->periodic ? bert_source_kernel<true, false> : bert_source_kernel<false, false>),
                               dim3(b->src_blocks),
                               dim3(256), b->lds_bytes, s, P, b->d_src.p, b->nsrc, b->d_src_tab.p, (uint32_t)istep,
                               cur, hist);
        hipLaunchKernelGGL(bert_wallcdf_kernel, dim3((unsigned)b->g->sides.size()), dim3(256), 0, s, P, (uint32_t)istep);
        hipLaunchKernelGGL(emit_k, dim3(emit_blocks), dim3(b->vl ? 1024 : BERT_EMIT_THREADS),
                           b->vl ? b->lds_vl_bytes : b->lds_bytes, s, P, (uint32_t)istep, cur, hist);
        // H of this step lives in H[cur]; H[cur^1] (previous step) is cleared
        // after the step kernel used it for cancellation
        hipLaunchKernelGGL(bert_props_kernel, dim3(nbox), dim3(256), 0, s, P, istep + 1, cur, cancel_now,
                           cancel_now, (istep + 1) & 1, P.lds_props, prev_hist);
        prev_hist = hist;
        HIPCHECK(hipGetLastError());
        if (cb) {
            // the histogram of this step is in H[cur] until the next step's
            // props kernel clears it
            HIPCHECK(hipMemcpyAsync(b->hH64.data(), b->d_H.p + (size_t)cur * nbox * n,
                                    (size_t)nbox * n * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            for (size_t k = 0; k < b->hH.size(); k++)
                b->hH[k] = (int32_t)(uint32_t)b->hH64[k] - (int32_t)(uint32_t)(b->hH64[k] >> 32);
            if (cb(user, istep, (istep + 1) * b->cfg.dt, b->Eeff, nbox, n, b->hH.data())) {
                istep++;
                break;
            }
        }
    }
    HIPCHECK(hipEventRecord(b->ev1, s));
    HIPCHECK(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    double kms = 0;
    int ktimed = 0;
    for (int i = 0; i < istep; i++) {
        if ((i % pc) != ((i / pc) % pc)) continue;
        float e = 0;
        HIPCHECK(hipEventElapsedTime(&e, b->kev[2 * i], b->kev[2 * i + 1]));
        kms += e;
        ktimed++;
    }
    uint32_t flags[4];
    unsigned long long stats[2];
    HIPCHECK(hipMemcpy(flags, b->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(stats, b->d_stats.p, sizeof(stats), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(props, b->d_props.p, (size_t)(istep + 1) * nbox * 3 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(npart, b->d_npart.p, (size_t)(istep + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (pending)
        HIPCHECK(hipMemcpy(pending, b->d_pending.p, (size_t)(istep + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (flags[0]) throw Error(BTE_ECAPACITY, "particle buffer capacity exceeded; raise capacity");
    if (flags[1]) throw Error(BTE_EINVAL, "a particle reached a mode outside the propagator");
    if (flags[2]) throw Error(BTE_EGEOM, "advection or branching exceeded its iteration bound");
    if (flags[3]) throw Error(BTE_EGEOM, "Error in translation, no box found");
    if (rep) {
        rep->nsteps = istep;
        rep->Eeff = b->Eeff;
        rep->particle_steps = stats[0] + stats[1];
        rep->step_kernel_particle_steps = stats[0];
        rep->step_kernel_launches = istep;
        int64_t mx = 0;
        for (int i = 0; i <= istep; i++) mx = std::max(mx, npart[i]);
        rep->max_particles = mx;
        rep->loop_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        rep->step_kernel_ms = kms;
        rep->step_kernel_timed = ktimed;
        rep->step_threads = b->vl ? 1024 : 256;
    }
    BTE_GUARD_END
}

int bte_bert_destroy(bte_bert* bb) {
    BTE_GUARD_BEGIN
    delete reinterpret_cast<Bert*>(bb);
    BTE_GUARD_END
}

}  // extern "C"
