/* C ABI of the MI355X deviational Monte Carlo + Krylov propagator engine.
 *
 * Plain pointers and sizes; no exceptions cross this boundary. Every entry
 * point returns BTE_OK (0) or a negative BTE_E* code; bte_last_error() gives
 * the thread's last message. Host arrays are caller-owned; device memory is
 * owned by the handles. Calls on one context are externally synchronised.
 *
 * The reference has no FFI: its boundary is the three executables and the
 * C++ interfaces they link statically (src/CMakeLists.txt:2-21). Each entry
 * point below names the reference interface it replaces.
 */
#ifndef BTE_ABI_H
#define BTE_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTE_OK 0
#define BTE_EINVAL -1    /* bad argument (reference: value_error/input_error) */
#define BTE_EGEOM -2     /* geometry_error (include/exceptions.hpp:33-54) */
#define BTE_EHIP -3      /* HIP runtime failure */
#define BTE_ENOMEM -4    /* device allocation failed */
#define BTE_ECAPACITY -5 /* particle buffer capacity exceeded */
#define BTE_ENODEV -6    /* no GPU visible */

typedef struct bte_ctx bte_ctx;
typedef struct bte_modes bte_modes;
typedef struct bte_geom bte_geom;
typedef struct bte_prop bte_prop;
typedef struct bte_rta bte_rta;
typedef struct bte_bert bte_bert;

/* ---- errors, version ---- */
int bte_last_error(char* buf, size_t len);
int bte_version(void);

/* ---- context: one HIP device + one stream ---- */
int bte_ctx_create(int device, bte_ctx** out);
int bte_ctx_destroy(bte_ctx* ctx);
int bte_sync(bte_ctx* ctx);

/* ---- host utilities (no GPU use) ---- */
/* Seeded synthetic mode table: f ~ U(fmin,fmax) THz, omega = 2 pi f,
 * |v| ~ U(vmin,vmax) nm/ps at angle ~ U(0,2pi), tau = 10^U(log tmin, log tmax)
 * ps, w = 1/tau; mirror != 0 repeats the first half with vy -> -vy. Stands in
 * for the Gamma_grid loaded by load_bulk_hdf5 (src/bulk_hdf5.cpp:443). */
int bte_synthetic_modes(int n, uint64_t seed, double fmin, double fmax, double vmin,
                        double vmax, double tmin, double tmax, int mirror, double* omega,
                        double* vx, double* vy, double* w);
/* Synthetic collision operator B = -Omega (N x N, column-major): symmetric
 * nonnegative pattern of the given density scaled so that 1^T B = 0 and
 * B C = 0. Stands in for get_collision_operator_dense
 * (src/collision_operator.cpp:38-812). */
int bte_synthetic_B(int n, const double* C, const double* w, double density, uint64_t seed,
                    int sinkhorn_iters, double* B);
/* Vose alias table of a nonnegative pmf: thresh in units of 2^-32. */
int bte_alias_build(int n, const double* pmf, uint32_t* thresh, uint32_t* alias);

/* ---- material: phonon modes (replaces alma::Gamma_grid + Crystal_structure
 * as used by the MC solvers: omega, vg, V, lattvec(2,2);
 * include/dynamical_matrix.hpp:88-95, include/structures.hpp) ---- */
int bte_modes_create(bte_ctx* ctx, int n, int nq, const double* omega, const double* vx,
                     const double* vy, const double* w, double V, double lattz,
                     bte_modes** out);
/* Per-mode broadening sigma[n] [rad/ps] of the diffuse bounce: base_sigma
 * of a real q-grid, ||v^T dq|| / sqrt(12) (qpoint_grid.hpp:179-181), given
 * with the mode table since this engine reads no q-grid. Lambert tables of
 * RTAMC2D runs created afterwards use hypot(sigma_i, sigma_f) per pair after
 * the reference's low-outlier floor (RTA_MC2D.cpp:566-618) instead of the
 * constant sigma_bounce; spectral broadening without an explicit sigma uses
 * it too. NULL clears it. */
int bte_modes_set_sigma(bte_modes* m, const double* sigma);
int bte_modes_destroy(bte_modes* m);

/* ---- geometry: axis-aligned rectangular boxes (replaces
 * alma::read_geometry_XML -> std::vector<alma::geometry_2d>,
 * src/geometry_2d.cpp:408-522). rect: xmin,xmax,ymin,ymax per box.
 * kind: 0 regular box, 1 reservoir (<Reservoir/>), 2 periodic (then
 * bte_geom_set_periodic gives its translation).
 * Tinit may be NULL (Tinit = Teq). ---- */
int bte_geom_create_rects(bte_ctx* ctx, int nbox, const double* rect, const int32_t* kind,
                          const double* Teq, const double* Tinit, bte_geom** out);
/* Convex polygon boxes (read_geometry_XML with <Vertices npoints=n>,
 * geometry_2d.cpp:127-153 and 408-522): box i has nvert[i] vertices, its
 * (x, y) pairs consecutive in xy. Each box is the convex hull of its
 * vertices (at most 16 sides); a vertex on a side between two hull corners
 * is refused (the reference keeps only part of such a side as a border).
 * Contacts are the boundary segments or points two boxes share
 * (geometry_2d.cpp:295-381); overlapping boxes are an error. Runs on such a
 * geometry use the polygon kernels (diffuse sides; specular sides and
 * dT/dp need rectangles). */
int bte_geom_create_polys(bte_ctx* ctx, int nbox, const int32_t* nvert, const double* xy, const int32_t* kind,
                          const double* Teq, const double* Tinit, bte_geom** out);
int bte_geom_destroy(bte_geom* g);
/* Periodic box (kind 2, <Translate_to id> tx ty </Translate_to>,
 * geometry_2d.cpp:465-473): a particle entering it is moved by 1.0005 (tx, ty)
 * into the regular neighbour of box `target` (geometry_2d.hpp:366-494). */
int bte_geom_set_periodic(bte_geom* g, int box, int target, double tx, double ty);
/* Contacts (geometry_2d::calculate_contacts, src/geometry_2d.cpp:295-381):
 * for each box up to maxc neighbours with (npx, npy, p1x, p1y, p2x, p2y). */
int bte_geom_contacts(const bte_geom* g, int maxc, int32_t* ncont, int32_t* nbr, double* geo);
/* Temperature gradients (calculate_gradientT, src/geometry_2d.cpp:524-576):
 * per box (gx, gy) [K/nm] into grad[2*nbox], the minimum-norm least-squares
 * fit of the Teq differences to the regular contact boxes; 0 for reservoir
 * and periodic boxes. Host only. */
int bte_geom_gradient_T(const bte_geom* g, double* grad);

/* ---- RTAMC2D transient solver (replaces the particle loop of
 * src/RTA_MC2D.cpp:2188-2435: Init_generator, evolparticle, process_segment,
 * diffusive_bouncing and the post-processing to temperature/jx/jy) ---- */
typedef struct {
    double thickness;    /* <material thickness> [nm] */
    double T0;           /* <material T0> [K] */
    double dt, maxtime;  /* <time dt maxtime> [ps] */
    uint64_t nparticles; /* <particles N> */
    uint64_t seed;
    int ballistic;       /* <ballistic/> */
    double sigma_bounce; /* Lambert broadening per mode [rad/ps] */
} bte_rta_cfg;

typedef struct {
    uint64_t events;    /* scatters + wall bounces + reservoir absorptions */
    uint64_t crossings; /* box-to-box crossings */
    double Eeff;        /* particle power [J/ps] */
    double kernel_ms;   /* device time of the history kernel */
    int ntime;          /* number of time rows written (ntime-1 bins) */
} bte_rta_report;

/* Outputs are [ntime-1][nbox] (the CSV rows of temperature_/jx_/jy_ without
 * the time column, RTA_MC2D.cpp:2437-2471). */
int bte_rta_create(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_rta_cfg* cfg,
                   bte_rta** out);
/* bte_rta_execute switches the handle to tally mode 2 (FP64 LDS sums of
 * the leading time bins, one 768-thread workgroup per CU) whenever the
 * run's cells fit and no spectral location is set, and back to mode 0
 * (256-thread workgroups, eight per CU) otherwise. */
int bte_rta_execute(bte_rta* r, double* T_out, double* jx_out, double* jy_out,
                    bte_rta_report* rep);
/* No counterpart in the reference: bte_rta_execute for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops). The same run (same streams, tally mode and post-processing) with
 * its tally copies, counters and per-workgroup blocks in the library's
 * static device scratch array of 64 MiB per GPU instead of buffers it
 * allocates; runs whose tallies need more fail with BTE_EINVAL before any
 * launch. Outputs are device pointers: T, jx, jy [(ntime-1) x nbox] doubles
 * as bte_rta_execute writes them, counts[6] uint64 (events, crossings, ids
 * taken, stuck histories, missing mirror partners, failed translations; the
 * last three are errors of the run). Asynchronous on the context's stream;
 * it makes no device allocation and no host synchronisation. Single-run
 * handles without spectral locations or gradient sides only. Calls that
 * overlap on one GPU share the scratch array and race: serialise them (one
 * stream, or the caller's own ordering). */
int bte_rta_execute_noalloc(bte_rta* r, double* T, double* jx, double* jy, unsigned long long* counts);
/* Spectral decomposition (<spectral><resolution ticks/><location bin/>,
 * RTA_MC2D.cpp:1930-1945, 1978-2052; process_segment 718-766; outputs
 * 2349-2405, 2473-2541), set before bte_rta_execute: the segments in the nloc
 * boxes are also tallied per mode and broadened afterwards with Gamma
 * densities of shape omega^2/sigma and scale sigma/omega on nomega
 * frequency ticks. sigma: per-mode broadening [rad/ps]
 * (Gamma_grid::base_sigma, qpoint_grid.hpp:179-181), or NULL for the run's
 * sigma_bounce; low outliers are raised to exp(p25 - 1.5 (p75 - p25)) of
 * log sigma. nloc = 0 switches it off. */
int bte_rta_spectral(bte_rta* r, int nloc, const int32_t* boxes, int nomega, const double* sigma);
/* After bte_rta_execute, location k: omega[nomega] tick centres,
 * dT[(ntime-1) x nomega] (deltaT_omega), jx, jy [(ntime-1) x nomega]
 * (jx_omega, jy_omega), fd_q [(ntime-1) x nq] (fd_q); rows are the
 * cumulative time bins like the temperature output. Outputs may be NULL. */
int bte_rta_spectral_get(const bte_rta* r, int k, double* omega, double* dT, double* jx, double* jy,
                         double* fd_q);
int bte_rta_destroy(bte_rta* r);
/* Gamma-broadened spectra of per-mode weights (dist_reader's
 * print_distribution, dist_reader.cpp:290-380, on a frame's histogram; the
 * same broadening as the RTAMC2D spectra): out[nrows][3 nomega + nq] =
 * X[nrows][n] x W with W = (gamma_pdf(omega^2/sigma, sigma/omega, grid_j),
 * vx times it, vy times it, 1/omega on the mode's q column). sigma: per
 * mode, or NULL for sigma_default; low outliers raised as in
 * bte_rta_spectral. One matrix product on the GPU. */
int bte_spectra(bte_ctx* ctx, const bte_modes* m, const double* sigma, double sigma_default, int nomega,
                const double* grid, int nrows, const double* X, double* out);
int bte_rta_run(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_rta_cfg* cfg,
                double* T_out, double* jx_out, double* jy_out, bte_rta_report* rep);

/* ---- RTAMC2D steady branch (a nonzero <gradient>; replaces the
 * generation loop of src/RTA_MC2D.cpp:2544-2782 with Nabla_T sources,
 * RTA_MC2D.cpp:347-533, and the steady evolparticle, 1406-1799) ---- */
typedef struct {
    double thickness, T0;
    double gradient[2];      /* <gradient x y> [K/nm] */
    uint64_t nparticles, seed;
    double sigma_bounce;
    double eps_energy, eps_flux; /* <convergence energy flux>; <= 0: until no particle is left */
    int max_generations;     /* 0: no cap */
} bte_rta_steady_cfg;

typedef struct {
    int generations;
    int converged;
    uint64_t flights;
    double Eeff;
    double kernel_ms;        /* device time of the flight kernels */
} bte_rta_steady_report;

/* Outputs per box (the rows of steady_<T0>K_run_<i>.csv): T [K], jx, jy [W/m^2]. */
int bte_rta_steady_run(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_rta_steady_cfg* cfg,
                       double* T, double* jx, double* jy, bte_rta_steady_report* rep);

/* ---- batched RTAMC2D: independent runs of one material (geometry,
 * temperature, particle count, seed and edge specularity per run) in one
 * persistent launch. The reference runs its nruns one after another
 * (RTA_MC2D.cpp:2188-2471); batching, specular edges and the specularity
 * gradient have no counterpart there. ---- */
typedef struct bte_rta_batch bte_rta_batch;
typedef struct {
    const bte_geom* geom;   /* the run's boxes; regular boxes at Teq = T0 */
    double T0;              /* run temperature [K]: Cv(T0), BE_derivative at T0 */
    uint64_t nparticles;
    uint64_t seed;
    int ballistic;
    double wscale;          /* scattering rates used: w * wscale */
    double specularity[4];  /* probability of a specular reflection at a void
                               side (left, bottom, top, right); else diffuse */
    int grad;               /* accumulate dT/dp for each side with 0 < p < 1 */
} bte_rta_run_desc;

typedef struct {
    uint64_t events, crossings;
    double kernel_ms;       /* device time of the batch */
    int ntime;
    int block, blocks;      /* launch shape in use */
} bte_rta_batch_report;

/* use_graph: capture the clear + history launches once in a HIP graph and
 * replay it on every execute. */
int bte_rta_batch_create(bte_ctx* ctx, const bte_modes* m, int nruns, const bte_rta_run_desc* runs,
                         double thickness, double dt, double maxtime, double sigma_bounce, int use_graph,
                         bte_rta_batch** out);
/* Tally arithmetic (no counterpart in the reference, which sums FP64
 * thread-local arrays, RTA_MC2D.cpp:699-769 and 2326-2335). mode 0 (default):
 * FP64 device atomics. mode 1: every workgroup sums its segments in FP32
 * arrays in LDS (T, jx, jy and each differentiated side over the regular
 * boxes' cells) and adds them into the FP64 tallies once at the end; one
 * 768-thread workgroup per CU. BTE_EINVAL when those arrays exceed the
 * device's LDS per workgroup. mode 2: FP64 sums of T, jx and jy in LDS for
 * as many leading time bins of all regular boxes as fit, later bins by FP64
 * device atomics. mode 3: as mode 2 with each workgroup working through
 * items of one run's particles at a time and its LDS holding only that
 * run's boxes (added into the tallies after each item), so a batch of many
 * runs keeps as many bins in LDS as one run. Modes 2 and 3: FP64 sums
 * throughout; not with dT/dp or spectral locations (BTE_EINVAL). */
int bte_rta_batch_set_tally(bte_rta_batch* b, int mode);
/* Times the batch scaled to `fraction` of each run's particles over a set of
 * launch shapes and keeps the fastest (no-op under tally modes 1, 2 and 3). */
int bte_rta_batch_tune(bte_rta_batch* b, double fraction, int* block, int* blocks);
/* T, jx, jy: per run r, [ntime-1][nbox_r] blocks one after another; dTdp (may
 * be NULL): per run [4 sides][ntime-1][nbox_r]; Eeff (may be NULL): [nruns]. */
int bte_rta_batch_execute(bte_rta_batch* b, double* T, double* jx, double* jy, double* dTdp, double* Eeff,
                          bte_rta_batch_report* rep);
int bte_rta_batch_destroy(bte_rta_batch* b);

/* ---- propagator: dense P (ncol x ncol, column-major) for modes
 * [ngamma, ngamma+ncol) (replaces load_P + propagator_H per column,
 * src/collision_operator.cpp:854-1013, beyondRTA_MC2d.cpp:1356-1378).
 * A box scatters with the propagator tabulated at its Teq; T <= 0 makes a
 * single propagator apply at every box temperature. ---- */
int bte_prop_create(bte_ctx* ctx, int ncol, const double* P, int ngamma, double T,
                    bte_prop** out);
/* Propagators of one material at ntemp temperatures Ts (the <propagator T
 * file/> children of <material>, beyondRTA_MC2d.cpp:1352-1378): Ps holds
 * ntemp column-major ncol x ncol matrices. A box whose Teq is not tabulated
 * scatters with the lirp of two of them (collision_operator.cpp:918-929),
 * chosen by bte_bert_cfg.prop_interp. */
int bte_prop_create_multi(bte_ctx* ctx, int ncol, int ntemp, const double* Ps, const double* Ts,
                          int ngamma, bte_prop** out);
int bte_prop_destroy(bte_prop* p);
/* The distribution of propagator_H::get(R) (collision_operator.cpp:892-904)
 * for R uniform in [0, 1) over a signed cumulative (propagator_H's
 * signed_cumm, possibly interpolated and then not monotone in |.|): pmf[k]
 * = probability of picking k; a search past the end picks n-1. Host only. */
int bte_prop_pick_pmf(int n, const double* signed_cumm, double* pmf);

/* ---- beRTAMC2D (replaces the step loop of src/beyondRTA_MC2d.cpp:
 * 2138-2425: advection, adiabatic/isothermal wall generators, Scattering or
 * ScatteringRTA, histogram, cancellation, calculate_prop, frame saving) ---- */
typedef struct {
    double thickness;  /* [nm] */
    double dt, maxtime;
    int max_steps;     /* cap on steps (<=0: run while time <= maxtime) */
    double Ed;         /* <Eeff Ed>; < 0 to derive from Efactor/Tmax/Tmin */
    uint64_t Efactor;  /* <Eeff particles> */
    double Tmax, Tmin; /* <Eeff Tmax Tmin>, 0 if absent */
    int rta;           /* <RTA/> */
    int pcancel_each;  /* cancellation period (4 in the reference) */
    uint64_t seed;
    uint64_t capacity; /* particle buffer capacity; 0 = automatic */
    /* gradient sources (source_term_gradient, beyondRTA_MC2d.cpp:597-676):
     * 0 = per-box gradients from calculate_gradientT (the default),
     * 1 = <gradient x y> [K/nm] in every regular box */
    int fixed_gradient;
    double gradient[2];
    /* pair of tabulated propagators for an untabulated box Teq:
     * 0 = Scattering's choice (beyondRTA_MC2d.cpp:1139-1164: the first
     *     temperature not below Teq and the next one, extrapolated down),
     * 1 = the adjacent temperatures bracketing Teq */
    int prop_interp;
    /* <init_distribution box file> (beyondRTA_MC2d.cpp:1411-1417, sampled
     * as init_particles 142-193 with ref_distribution, sampling.cpp:70-114):
     * ninit regular boxes init_boxes[ninit], each with a signed distribution
     * over the modes after the propagator's Gamma offset,
     * init_dist[ninit][init_len] (init_len = modes - ngamma); 0 / NULL for
     * none. These particles are kept together with those of boxes at
     * Tinit != Teq (the reference's later init_particles clears them,
     * beyondRTA_MC2d.cpp:2096-2105). */
    int ninit;
    int init_len;
    const int32_t* init_boxes;
    const double* init_dist;
    /* launch shape of the step and emission kernels (no counterpart in the
     * reference; the histories are the same either way): 0 = automatic,
     * the mode velocities and their inverses staged in LDS (one
     * 1024-thread workgroup per CU) when no box is periodic and 32 B per
     * mode fit next to the particle kernels' LDS layout, else 1; 1 = 256-
     * (step) and 512-thread (emission) workgroups reading them from global
     * memory */
    int step_shape;
} bte_bert_cfg;

typedef struct {
    int nsteps;
    double Eeff;
    uint64_t particle_steps;  /* particles advected and scattered, all kernels */
    uint64_t step_kernel_particle_steps; /* the part done by the main step kernel */
    int64_t max_particles;
    double loop_ms;           /* wall time of the step loop (device-synchronised) */
    double step_kernel_ms;    /* device time of the timed step-kernel launches, summed */
    int step_kernel_launches;
    /* launches bracketed by events: one in pcancel_each, rotating through
     * the cancellation cycle so every step type is sampled equally (an
     * event pair costs about 11 us of idle device time per step) */
    int step_kernel_timed;
    int step_threads;         /* threads per step-kernel workgroup: 1024 with the
                                 mode velocities in LDS, else 256 */
} bte_bert_report;

/* Frame callback: called after each step with the (box, mode) histogram
 * [nbox][nmodes] (properties.msgpack.bin, beyondRTA_MC2d.cpp:1604-1769).
 * Return nonzero to stop. */
typedef int (*bte_frame_cb)(void* user, int istep, double time, double Eeff, int nbox,
                            int nmodes, const int32_t* H);

/* props: [nsteps+1][nbox][3] (ed J/m^3, jx, jy W/m^2); npart: [nsteps+1]
 * (the stdout columns of beyondRTA_MC2d.cpp:2418-2423); pending (may be
 * NULL): [nsteps+1] signed wall counts (Eborder) waiting for re-emission at
 * the end of each step, so that npart-weighted signs + pending is conserved
 * with adiabatic walls and a nonnegative propagator. Capacity: max_steps+1. */
int bte_bert_create(bte_ctx* ctx, const bte_modes* m, const bte_geom* g, const bte_prop* p,
                    const bte_bert_cfg* cfg, bte_bert** out);
int bte_bert_execute(bte_bert* b, double* props, int64_t* npart, int64_t* pending, bte_frame_cb cb,
                     void* user, bte_bert_report* rep);
int bte_bert_destroy(bte_bert* b);

/* ---- Krylov propagator (replaces build_P + Arnoldi_algorithm,
 * src/collision_operator.cpp:815-851, 1015-1064): P = exp(dt*B), all columns
 * at once as block-Krylov GEMMs. ---- */
#define BTE_PREC_FP64 0
#define BTE_PREC_FP32 1
#define BTE_PREC_BF16 2
#define BTE_ARNOLDI_REFERENCE 0 /* projections over j<k only, as collision_operator.cpp:831-850 */
#define BTE_ARNOLDI_TEXTBOOK 1  /* modified Gram-Schmidt over j<=k */
typedef struct {
    double build_ms;   /* device time of the whole build */
    double gemm_ms;    /* device time in the block GEMMs */
    double gemm_tflops; /* algorithmic: 2 N^3 per step over gemm_ms, whatever the
                           number of partial products (BF16 forms each product
                           from three) */
    double max_col_sum_err; /* max_j |sum_i P_ij - 1| after the build */
    int breakdowns;    /* columns whose Krylov space closed early */
} bte_krylov_report;

int bte_krylov_build(bte_ctx* ctx, int N, const double* B, double dt, int m, int precision,
                     int arnoldi_mode, double* P, bte_krylov_report* rep);

/* Krylov size rule of build_P (collision_operator.cpp:1020-1030 with
 * error_bound, collision_operator.hpp:216-219): the smallest m >= 2 with
 * 2 |dt B|_F^m e^{|dt B|_F} / Gamma(m) / N^2 < eps. Host only. */
int bte_krylov_size(int N, const double* B, double dt, double eps, int* m);

/* FP64 exponential E = exp(t A) of the whole matrix (N x N, column-major):
 * Pade(13) scaling and squaring (Higham 2005) with FP64 GEMMs and an LU
 * solve on the GPU. No counterpart in the reference: it checks propagators
 * and replaces build_P (collision_operator.cpp:1015-1064) when every column
 * is wanted. ms (may be NULL): device time. */
int bte_expm(bte_ctx* ctx, int N, const double* A, double t, double* E, double* ms);

#ifdef __cplusplus
}
#endif
#endif
