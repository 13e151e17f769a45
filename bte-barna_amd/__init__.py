"""MI355X engine for BTE-Barna's deviational Monte Carlo (RTAMC2D,
beRTAMC2D) and its Krylov propagator builder, driven through the C ABI of
libbte_amd.so.

Python here is a thin host mirror used by the tests and the benchmark; the
C++ front-ends in apps/ are the drop-in executables.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (BertCfg, BertReport, BteError, FRAME_CB, KrylovReport, PREC, ARNOLDI, RtaCfg,
                   RtaReport, RtaRunDesc, RtaBatchReport, RtaSteadyCfg, RtaSteadyReport, check, lib, LIBPATH)
from . import workloads  # noqa: F401

__all__ = ["Context", "ModeTable", "Geometry", "Propagator", "rtamc2d", "bertamc2d", "build_P",
           "synthetic_modes", "synthetic_B", "alias_build", "BteError", "workloads", "LIBPATH"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def synthetic_modes(n, seed, fmin=0.1, fmax=15.0, vmin=0.5, vmax=8.0, tmin=1.0, tmax=1000.0,
                    mirror=False):
    """Seeded synthetic mode table -> dict(omega, vx, vy, w)."""
    out = {k: np.zeros(n) for k in ("omega", "vx", "vy", "w")}
    check(lib().bte_synthetic_modes(n, seed, fmin, fmax, vmin, vmax, tmin, tmax, int(mirror),
                                    _dp(out["omega"]), _dp(out["vx"]), _dp(out["vy"]), _dp(out["w"])))
    return out


def synthetic_B(C_modes, w, density=0.05, seed=1, iters=200):
    """Synthetic energy-conserving collision operator B = -Omega (column-major)."""
    C_modes, w = _f64(C_modes), _f64(w)
    n = len(C_modes)
    B = np.zeros((n, n), order="F")
    check(lib().bte_synthetic_B(n, _dp(C_modes), _dp(w), density, seed, iters,
                                B.ctypes.data_as(C.POINTER(C.c_double))))
    return B


def alias_build(pmf):
    pmf = _f64(pmf)
    n = len(pmf)
    th = np.zeros(n, np.uint32)
    al = np.zeros(n, np.uint32)
    check(lib().bte_alias_build(n, _dp(pmf), th.ctypes.data_as(C.POINTER(C.c_uint32)),
                                al.ctypes.data_as(C.POINTER(C.c_uint32))))
    return th, al


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = C.c_void_p()

    def close(self):
        if self.h and self.h.value:
            check(getattr(lib(), self._destroy)(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context(_Handle):
    _destroy = "bte_ctx_destroy"

    def __init__(self, device=0):
        super().__init__()
        check(lib().bte_ctx_create(device, C.byref(self.h)))

    def sync(self):
        check(lib().bte_sync(self.h))


class ModeTable(_Handle):
    """Phonon modes of one material: omega [rad/ps], v [nm/ps], w [1/ps],
    unit-cell volume V [nm^3], lattvec(2,2) [nm], nq q-points."""
    _destroy = "bte_modes_destroy"

    def __init__(self, ctx, omega, vx, vy, w, V, lattz=1.0, nq=1):
        super().__init__()
        self.ctx = ctx
        self.omega, self.vx, self.vy, self.w = map(_f64, (omega, vx, vy, w))
        self.V, self.lattz, self.nq = float(V), float(lattz), int(nq)
        self.n = len(self.omega)
        check(lib().bte_modes_create(ctx.h, self.n, self.nq, _dp(self.omega), _dp(self.vx), _dp(self.vy),
                                     _dp(self.w), self.V, self.lattz, C.byref(self.h)))
        self.sigma = None

    def set_sigma(self, sigma):
        """per-mode broadening [rad/ps] of the diffuse bounce (base_sigma);
        None clears it (runs then use their constant sigma_bounce)"""
        self.sigma = None if sigma is None else _f64(sigma)
        check(lib().bte_modes_set_sigma(self.h, None if self.sigma is None else _dp(self.sigma)))


class Geometry(_Handle):
    """Rectangular boxes: rect rows (xmin, xmax, ymin, ymax) [nm]; or convex
    polygons: polys, a list of (n_i, 2) vertex arrays [nm] (rect is then
    ignored). kind 0 box, 1 reservoir, 2 periodic; periodic = {box:
    (target, tx, ty)} gives each periodic box its translation."""
    _destroy = "bte_geom_destroy"

    def __init__(self, ctx, rect, kind, Teq, Tinit=None, periodic=None, polys=None):
        super().__init__()
        self.ctx = ctx
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.Teq = _f64(Teq)
        self.Tinit = None if Tinit is None else _f64(Tinit)
        self.nbox = len(self.kind)
        tin = None if self.Tinit is None else _dp(self.Tinit)
        kp = self.kind.ctypes.data_as(C.POINTER(C.c_int32))
        if polys is not None:
            if len(polys) != self.nbox:
                raise ValueError("one vertex array per box")
            self.polys = [_f64(p).reshape(-1, 2) for p in polys]
            self.nvert = np.array([len(p) for p in self.polys], np.int32)
            self.xy = _f64(np.concatenate(self.polys))
            self.rect = np.array([[p[:, 0].min(), p[:, 0].max(), p[:, 1].min(), p[:, 1].max()] for p in self.polys])
            check(lib().bte_geom_create_polys(ctx.h, self.nbox, self.nvert.ctypes.data_as(C.POINTER(C.c_int32)),
                                              _dp(self.xy), kp, _dp(self.Teq), tin, C.byref(self.h)))
        else:
            self.polys = None
            self.rect = _f64(rect).reshape(-1, 4)
            check(lib().bte_geom_create_rects(ctx.h, self.nbox, _dp(self.rect), kp, _dp(self.Teq), tin,
                                              C.byref(self.h)))
        for b, (t, tx, ty) in (periodic or {}).items():
            check(lib().bte_geom_set_periodic(self.h, int(b), int(t), float(tx), float(ty)))

    def contacts(self, maxc=8):
        nc = np.zeros(self.nbox, np.int32)
        nbr = np.zeros((self.nbox, maxc), np.int32)
        geo = np.zeros((self.nbox, maxc, 6))
        check(lib().bte_geom_contacts(self.h, maxc, nc.ctypes.data_as(C.POINTER(C.c_int32)),
                                      nbr.ctypes.data_as(C.POINTER(C.c_int32)), _dp(geo)))
        return nc, nbr, geo

    def gradient_T(self):
        """calculate_gradientT: (nbox, 2) gradients [K/nm] used by the
        gradient sources"""
        g = np.zeros((self.nbox, 2))
        check(lib().bte_geom_gradient_T(self.h, _dp(g)))
        return g


class Propagator(_Handle):
    """Dense propagator P (ncol x ncol) acting on modes [ngamma, ngamma+ncol).
    P may be a list of matrices tabulated at temperatures T (a list); a
    single matrix with T = 0 applies at every box temperature."""
    _destroy = "bte_prop_destroy"

    def __init__(self, ctx, P, ngamma=0, T=0.0):
        super().__init__()
        Ps = list(P) if isinstance(P, (list, tuple)) else [P]
        Ts = np.atleast_1d(np.asarray(T, dtype=np.float64))
        if len(Ts) != len(Ps):
            raise ValueError("one temperature per propagator")
        self.ncol = np.asarray(Ps[0]).shape[0]
        flat = np.concatenate([np.asfortranarray(x, dtype=np.float64).ravel(order="F") for x in Ps])
        check(lib().bte_prop_create_multi(ctx.h, self.ncol, len(Ps), _dp(flat), _dp(Ts), ngamma, C.byref(self.h)))


def prop_pick_pmf(signed_cumm):
    """Probabilities of propagator_H::get over a signed cumulative (host only)."""
    sc = np.ascontiguousarray(signed_cumm, dtype=np.float64)
    pmf = np.zeros(len(sc))
    check(lib().bte_prop_pick_pmf(len(sc), _dp(sc), _dp(pmf)))
    return pmf


class RtaRun(_Handle):
    """RTAMC2D session: device tables are built once, execute() runs the
    particle histories (the timed hot path)."""
    _destroy = "bte_rta_destroy"

    def __init__(self, ctx, modes, geom, thickness, T0, dt, maxtime, nparticles, seed=1,
                 ballistic=False, sigma_bounce=0.25):
        super().__init__()
        self.cfg = RtaCfg(thickness, T0, dt, maxtime, nparticles, seed, int(ballistic), sigma_bounce)
        self.nbox = geom.nbox
        self.ntime = int(np.floor(maxtime / dt + 1e-6)) + 1
        self._keep = (ctx, modes, geom)
        check(lib().bte_rta_create(ctx.h, modes.h, geom.h, C.byref(self.cfg), C.byref(self.h)))
        self.spectral_boxes, self.nomega, self.nq = [], 0, modes.nq

    def spectral(self, boxes, nomega, sigma=None):
        """<spectral>: per-mode tallies in these boxes, broadened on nomega
        ticks after each run (sigma: per-mode broadening or None)"""
        b = np.ascontiguousarray(boxes, np.int32)
        sg = None if sigma is None else _dp(_f64(sigma))
        check(lib().bte_rta_spectral(self.h, len(b), b.ctypes.data_as(C.POINTER(C.c_int32)), int(nomega), sg))
        self.spectral_boxes, self.nomega = list(b), int(nomega)

    def execute_noalloc(self, T_ptr, jx_ptr, jy_ptr, counts_ptr):
        """bte_rta_execute_noalloc: the same run into device buffers (T, jx,
        jy: (ntime-1) x nbox doubles; counts: 6 uint64 = events, crossings,
        ids taken, stuck, no mirror, no translation), asynchronous on the
        context's stream, with no device allocation. Pointers are device
        addresses (ints)."""
        check(lib().bte_rta_execute_noalloc(self.h, C.c_void_p(T_ptr), C.c_void_p(jx_ptr), C.c_void_p(jy_ptr),
                                            C.c_void_p(counts_ptr)))

    def execute(self):
        nt = self.ntime - 1
        T = np.zeros((nt, self.nbox))
        jx = np.zeros((nt, self.nbox))
        jy = np.zeros((nt, self.nbox))
        rep = RtaReport()
        check(lib().bte_rta_execute(self.h, _dp(T), _dp(jx), _dp(jy), C.byref(rep)))
        out = {"T": T, "jx": jx, "jy": jy, "events": rep.events, "crossings": rep.crossings,
               "Eeff": rep.Eeff, "kernel_ms": rep.kernel_ms}
        if self.spectral_boxes:
            out["spectral"] = []
            for k in range(len(self.spectral_boxes)):
                d = {"omega": np.zeros(self.nomega), "dT": np.zeros((nt, self.nomega)),
                     "jx": np.zeros((nt, self.nomega)), "jy": np.zeros((nt, self.nomega)),
                     "fd": np.zeros((nt, self.nq))}
                check(lib().bte_rta_spectral_get(self.h, k, *(_dp(d[x]) for x in ("omega", "dT", "jx", "jy", "fd"))))
                out["spectral"].append(d)
        return out


def rtamc2d(ctx, modes, geom, spectral=None, **kw):
    """spectral: {"boxes": [...], "nomega": K, "sigma": None or per mode}"""
    r = RtaRun(ctx, modes, geom, **kw)
    try:
        if spectral:
            r.spectral(spectral["boxes"], spectral["nomega"], spectral.get("sigma"))
        return r.execute()
    finally:
        r.close()


def rta_steady(ctx, modes, geom, thickness, T0, gradient, nparticles, seed=1, sigma_bounce=0.25, eps_energy=-42.0,
               eps_flux=-42.0, max_generations=0):
    """RTAMC2D steady branch (nonzero gradient [K/nm]): per box T, jx, jy."""
    cfg = RtaSteadyCfg(thickness, T0, (C.c_double * 2)(*gradient), nparticles, seed, sigma_bounce, eps_energy,
                       eps_flux, max_generations)
    T, jx, jy = np.zeros(geom.nbox), np.zeros(geom.nbox), np.zeros(geom.nbox)
    rep = RtaSteadyReport()
    check(lib().bte_rta_steady_run(ctx.h, modes.h, geom.h, C.byref(cfg), _dp(T), _dp(jx), _dp(jy), C.byref(rep)))
    return {"T": T, "jx": jx, "jy": jy, "generations": rep.generations, "converged": bool(rep.converged),
            "flights": rep.flights, "Eeff": rep.Eeff, "kernel_ms": rep.kernel_ms}


class RtaBatch(_Handle):
    """Independent RTAMC2D runs of one material in one persistent launch.

    runs: list of dicts with geom, T0, nparticles, seed and optionally
    ballistic, wscale (scattering-rate factor), specularity (4 sides: left,
    bottom, top, right) and grad (dT/dp for sides with 0 < p < 1).

    tally: "fp64" (device atomics, the default), "fp32-lds" (per-workgroup
    FP32 partial sums in LDS added into the FP64 tallies at the end; raises
    when the regular boxes' cells do not fit in LDS), "fp64-lds" (FP64
    per-workgroup sums of T, jx and jy in LDS for the leading time bins that
    fit, later bins by device atomics; no dT/dp) or "fp64-lds-runs" (the same
    with each workgroup's LDS holding the rows of one run at a time: work
    items of one run's particles, so batches of many runs keep as many time
    bins in LDS as a single run; no dT/dp)."""
    _destroy = "bte_rta_batch_destroy"

    def __init__(self, ctx, modes, runs, thickness, dt, maxtime, sigma_bounce=0.25, use_graph=False,
                 tally="fp64"):
        super().__init__()
        self.nruns = len(runs)
        self.descs = (RtaRunDesc * self.nruns)()
        for d, r in zip(self.descs, runs):
            d.geom = r["geom"].h
            d.T0 = r["T0"]
            d.nparticles = int(r["nparticles"])
            d.seed = int(r.get("seed", 1))
            d.ballistic = int(r.get("ballistic", False))
            d.wscale = float(r.get("wscale", 1.0))
            sp = r.get("specularity", (0.0, 0.0, 0.0, 0.0))
            for k in range(4):
                d.specularity[k] = float(sp[k])
            d.grad = int(r.get("grad", False))
        self.nbox = [r["geom"].nbox for r in runs]
        self.ntime = int(np.floor(maxtime / dt + 1e-6)) + 1
        self._keep = (ctx, modes, [r["geom"] for r in runs])
        check(lib().bte_rta_batch_create(ctx.h, modes.h, self.nruns, self.descs, thickness, dt, maxtime,
                                         sigma_bounce, int(use_graph), C.byref(self.h)))
        modes_ = {"fp64": 0, "fp32-lds": 1, "fp64-lds": 2, "fp64-lds-runs": 3}
        if tally not in modes_:
            raise ValueError(f"tally must be one of {sorted(modes_)}")
        if tally != "fp64":
            check(lib().bte_rta_batch_set_tally(self.h, modes_[tally]))
        self.tally = tally

    def tune(self, fraction=0.05):
        blk, nb = C.c_int(0), C.c_int(0)
        check(lib().bte_rta_batch_tune(self.h, fraction, C.byref(blk), C.byref(nb)))
        return blk.value, nb.value

    def execute(self, grad=False):
        nt = self.ntime - 1
        tot = nt * sum(self.nbox)
        T, jx, jy = np.zeros(tot), np.zeros(tot), np.zeros(tot)
        dT = np.zeros(4 * tot) if grad else None
        E = np.zeros(self.nruns)
        rep = RtaBatchReport()
        check(lib().bte_rta_batch_execute(self.h, _dp(T), _dp(jx), _dp(jy), _dp(dT) if grad else None, _dp(E),
                                          C.byref(rep)))
        out, o = [], 0
        for nb in self.nbox:
            sl = slice(o, o + nt * nb)
            r = {"T": T[sl].reshape(nt, nb), "jx": jx[sl].reshape(nt, nb), "jy": jy[sl].reshape(nt, nb)}
            if grad:
                r["dTdp"] = dT[4 * o:4 * (o + nt * nb)].reshape(4, nt, nb)
            out.append(r)
            o += nt * nb
        return out, {"events": rep.events, "crossings": rep.crossings, "kernel_ms": rep.kernel_ms,
                     "block": rep.block, "blocks": rep.blocks, "Eeff": E}


class BertRun(_Handle):
    """beRTAMC2D session (device tables built once, execute() runs the steps)."""
    _destroy = "bte_bert_destroy"

    def __init__(self, ctx, modes, geom, prop, thickness, dt, maxtime, max_steps=0, Ed=-1.0,
                 Efactor=0, Tmax=0.0, Tmin=0.0, rta=False, pcancel_each=4, seed=1, capacity=0,
                 gradient=None, prop_interp=0, init_dist=None, step_shape=0):
        """gradient: (gx, gy) [K/nm] in every regular box (<gradient x y>);
        None derives per-box gradients from the box temperatures.
        init_dist: {box: signed distribution over the modes after the
        propagator's Gamma offset} (<init_distribution box file>).
        step_shape: 0 chooses the step kernel's launch shape (the mode
        velocities in LDS when they fit), 1 keeps them in global memory."""
        super().__init__()
        gx, gy = gradient if gradient is not None else (0.0, 0.0)
        self.cfg = BertCfg(thickness, dt, maxtime, max_steps, Ed, Efactor, Tmax, Tmin, int(rta),
                           pcancel_each, seed, capacity, int(gradient is not None), (C.c_double * 2)(gx, gy),
                           int(prop_interp))
        self.cfg.step_shape = int(step_shape)
        self._init = None
        if init_dist:
            boxes = sorted(init_dist)
            lens = {len(init_dist[b]) for b in boxes}
            if len(lens) != 1:
                raise ValueError("given distributions of different lengths")
            ib = np.ascontiguousarray(boxes, np.int32)
            fd = np.ascontiguousarray(np.concatenate([np.asarray(init_dist[b], np.float64) for b in boxes]))
            self._init = (ib, fd)
            self.cfg.ninit = len(boxes)
            self.cfg.init_len = lens.pop()
            self.cfg.init_boxes = ib.ctypes.data_as(C.POINTER(C.c_int32))
            self.cfg.init_dist = _dp(fd)
        self.nbox = geom.nbox
        nst = 0
        t = 0.0
        while t <= maxtime:
            nst += 1
            t += dt
        self.nsteps = min(max_steps, nst) if max_steps > 0 else nst
        self._keep = (ctx, modes, geom, prop)
        check(lib().bte_bert_create(ctx.h, modes.h, geom.h, None if prop is None else prop.h,
                                    C.byref(self.cfg), C.byref(self.h)))

    def execute(self, frame_cb=None):
        props = np.zeros((self.nsteps + 1, self.nbox, 3))
        npart = np.zeros(self.nsteps + 1, np.int64)
        pending = np.zeros(self.nsteps + 1, np.int64)
        rep = BertReport()
        cb = FRAME_CB(0) if frame_cb is None else FRAME_CB(
            lambda u, s, t, e, nb, nm, H: int(bool(frame_cb(s, t, e, np.ctypeslib.as_array(H, (nb, nm)).copy()))))
        check(lib().bte_bert_execute(self.h, _dp(props), npart.ctypes.data_as(C.POINTER(C.c_int64)),
                                     pending.ctypes.data_as(C.POINTER(C.c_int64)), cb, None, C.byref(rep)))
        n = rep.nsteps
        return {"props": props[:n + 1], "npart": npart[:n + 1], "pending": pending[:n + 1], "nsteps": n,
                "Eeff": rep.Eeff,
                "particle_steps": rep.particle_steps, "step_kernel_particle_steps": rep.step_kernel_particle_steps,
                "max_particles": rep.max_particles, "loop_ms": rep.loop_ms,
                "step_kernel_ms": rep.step_kernel_ms, "step_kernel_launches": rep.step_kernel_launches,
                "step_kernel_timed": rep.step_kernel_timed, "step_threads": rep.step_threads}


def bertamc2d(ctx, modes, geom, prop, **kw):
    b = BertRun(ctx, modes, geom, prop, **kw)
    try:
        return b.execute()
    finally:
        b.close()


def build_P(ctx, B, dt, m=30, precision="fp64", arnoldi="textbook"):
    """P = exp(dt*B) for all columns (block Krylov); returns (P, report)."""
    B = np.asfortranarray(B, dtype=np.float64)
    N = B.shape[0]
    P = np.zeros((N, N), order="F")
    rep = KrylovReport()
    check(lib().bte_krylov_build(ctx.h, N, B.ctypes.data_as(C.POINTER(C.c_double)), dt, m, PREC[precision],
                                 ARNOLDI[arnoldi], P.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rep)))
    return P, {"build_ms": rep.build_ms, "gemm_ms": rep.gemm_ms, "gemm_tflops": rep.gemm_tflops,
               "max_col_sum_err": rep.max_col_sum_err, "breakdowns": rep.breakdowns}


def expm(ctx, A, t=1.0):
    """exp(t A) in FP64 on the GPU (Pade(13) scaling and squaring); returns
    (E, device ms)"""
    A = np.asfortranarray(A, dtype=np.float64)
    E = np.zeros(A.shape, order="F")
    ms = C.c_double(0.0)
    check(lib().bte_expm(ctx.h, A.shape[0], A.ctypes.data_as(C.POINTER(C.c_double)), float(t),
                         E.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms)))
    return E, ms.value


def krylov_size(B, dt, eps=1e-8):
    """The reference's Krylov size rule for exp(dt*B) (host only)."""
    B = np.asfortranarray(B, dtype=np.float64)
    m = C.c_int(0)
    check(lib().bte_krylov_size(B.shape[0], B.ctypes.data_as(C.POINTER(C.c_double)), dt, eps, C.byref(m)))
    return m.value


APPS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


def app_path(name):
    """Path of a built front-end executable (RTAMC2D, beRTAMC2D, PropagatorBuilder)."""
    p = os.path.join(APPS_DIR, name)
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: run __graft_entry__.build() first")
    return p

