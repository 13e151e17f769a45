"""ctypes wrapper of the oracle (oracle/_build/liboracle.so). Test
infrastructure only: the oracle is the checker, never the thing measured."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
u64 = C.c_uint64

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run __graft_entry__.build()")
        _L = C.CDLL(ORACLE_SO)
        _L.orc_last_error.restype = C.c_char_p
        _L.orc_calc_cv.restype = C.c_double
        _L.orc_wall_dist.restype = C.c_double
        _L.orc_error_bound.restype = C.c_double
        _L.orc_gamma_pdf.restype = C.c_double
    return _L


def _dp(a):
    return a.ctypes.data_as(dp)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Pcg32:
    def __init__(self, state, stream):
        self.st = (C.c_uint64 * 2)()
        lib().orc_pcg32_init(u64(state), u64(stream), self.st)

    def next(self, n=1):
        out = (C.c_uint32 * n)()
        lib().orc_pcg32_next(self.st, n, out)
        return list(out)

    def bounded(self, bound, n=1):
        out = (C.c_uint32 * n)()
        lib().orc_pcg32_bounded(self.st, C.c_uint32(bound), n, out)
        return list(out)

    def uniform(self, n=1):
        out = np.zeros(n)
        lib().orc_pcg32_uniform(self.st, n, _dp(out))
        return out


def synthetic_modes(n, seed, fmin=0.1, fmax=15.0, vmin=0.5, vmax=8.0, tmin=1.0, tmax=1000.0, mirror=False):
    o = {k: np.zeros(n) for k in ("omega", "vx", "vy", "w")}
    lib().orc_synthetic_modes(n, u64(seed), C.c_double(fmin), C.c_double(fmax), C.c_double(vmin),
                              C.c_double(vmax), C.c_double(tmin), C.c_double(tmax), int(mirror),
                              _dp(o["omega"]), _dp(o["vx"]), _dp(o["vy"]), _dp(o["w"]))
    return o


def bose_einstein(omega, T):
    omega = _f(omega)
    out = np.zeros_like(omega)
    lib().orc_bose_einstein(len(omega), _dp(omega), C.c_double(T), _dp(out))
    return out


def bose_einstein_kernel(omega, T):
    omega = _f(omega)
    out = np.zeros_like(omega)
    lib().orc_bose_einstein_kernel(len(omega), _dp(omega), C.c_double(T), _dp(out))
    return out


def calc_cv(omega, V, T, nq=1):
    omega = _f(omega)
    return lib().orc_calc_cv(len(omega), _dp(omega), C.c_double(V), nq, C.c_double(T))


def wall_dist(omega, vx, vy, Twall, Teq, nx, ny, vuc, nq=1):
    omega, vx, vy = _f(omega), _f(vx), _f(vy)
    cum = np.zeros(len(omega))
    flux = C.c_double()
    cs = lib().orc_wall_dist(len(omega), _dp(omega), _dp(vx), _dp(vy), nq, C.c_double(Twall), C.c_double(Teq),
                             C.c_double(nx), C.c_double(ny), C.c_double(vuc), _dp(cum), C.byref(flux))
    return cs, cum, flux.value


def spline(X, Y, q):
    X, Y, q = _f(X), _f(Y), _f(q)
    out = np.zeros(len(q))
    lib().orc_spline(len(X), _dp(X), _dp(Y), len(q), _dp(q), _dp(out))
    return out


def grid_sample(p, seed, stream, ns):
    p = _f(p)
    out = np.zeros(ns, np.int32)
    lib().orc_grid_sample(len(p), _dp(p), u64(seed), u64(stream), ns, out.ctypes.data_as(ip32))
    return out


def propH_get(b, R):
    b, R = _f(b), _f(R)
    sg = np.zeros(len(R), np.int32)
    idx = np.zeros(len(R), np.int32)
    lib().orc_propH_get(len(b), _dp(b), len(R), _dp(R), sg.ctypes.data_as(ip32), idx.ctypes.data_as(ip32))
    return sg, idx


def propH_get_sc(sc, R):
    """propagator_H::get on a signed cumulative (e.g. an interpolated one)"""
    sc, R = _f(sc), _f(R)
    sg = np.zeros(len(R), np.int32)
    idx = np.zeros(len(R), np.int32)
    lib().orc_propH_get_sc(len(sc), _dp(sc), len(R), _dp(R), sg.ctypes.data_as(ip32), idx.ctypes.data_as(ip32))
    return sg, idx


def prop_interp(Ps, Ts, Tx, rule=0):
    """signed cumulatives (ncol x ncol, column j = propagator_H of column j)
    used at temperature Tx; raises ValueError when Tx cannot be interpolated"""
    order = np.argsort(Ts)
    Ts = _f(np.asarray(Ts, np.float64)[order])
    Pa = np.concatenate([np.asfortranarray(Ps[i], dtype=np.float64).ravel(order="F") for i in order])
    ncol = Ps[0].shape[0]
    sc = np.zeros((ncol, ncol), order="F")
    rc = lib().orc_prop_interp(ncol, len(Ts), _dp(Ts), _dp(Pa), C.c_double(Tx), rule, sc.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError(lib().orc_last_error().decode())
    return sc


def gradient_T(rect, kind, Teq, polys=None):
    """calculate_gradientT: (nbox, 2) gradients [K/nm]"""
    rect, Teq = _f(poly_rect(polys) if polys is not None else rect), _f(Teq)
    set_polygons(polys)
    kind = np.ascontiguousarray(kind, np.int32)
    g = np.zeros((len(kind), 2))
    rc = lib().orc_gradient_T(len(kind), _dp(rect), kind.ctypes.data_as(ip32), _dp(Teq), g.ctypes.data_as(dp))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return g


def inter_side(rect, ibox, x0, y0, vx, vy, dt, polys=None):
    rect = _f(poly_rect(polys) if polys is not None else rect)
    set_polygons(polys)
    out = np.zeros(3)
    bids = np.zeros(4, np.int32)
    nb = C.c_int32()
    rc = lib().orc_inter_side(len(rect) // 4 if rect.ndim == 1 else len(rect), _dp(rect), ibox, C.c_double(x0),
                              C.c_double(y0), C.c_double(vx), C.c_double(vy), C.c_double(dt), _dp(out),
                              bids.ctypes.data_as(ip32), C.byref(nb))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out[0], out[1], out[2], list(bids[:nb.value])


def contacts(rect, maxc=8, polys=None):
    rect = _f(poly_rect(polys) if polys is not None else rect)
    set_polygons(polys)
    nbox = len(rect)
    nc = np.zeros(nbox, np.int32)
    nbr = np.zeros((nbox, maxc), np.int32)
    geo = np.zeros((nbox, maxc, 6))
    rc = lib().orc_contacts(nbox, _dp(rect), maxc, nc.ctypes.data_as(ip32), nbr.ctypes.data_as(ip32), _dp(geo))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return nc, nbr, geo


def poly_rect(polys):
    """bounding boxes (xmin, xmax, ymin, ymax) of polygon vertex arrays"""
    return np.array([[np.min(p[:, 0]), np.max(p[:, 0]), np.min(p[:, 1]), np.max(p[:, 1])]
                     for p in (np.asarray(q, np.float64).reshape(-1, 2) for q in polys)])


def set_polygons(polys):
    """convex polygon boxes (list of (n_i, 2) vertex arrays, in the given
    vertex order) for the next oracle call in place of its rectangles"""
    if polys is None:
        return
    ps = [np.asarray(q, np.float64).reshape(-1, 2) for q in polys]
    nv = np.array([len(q) for q in ps], np.int32)
    xy = _f(np.concatenate(ps).ravel())
    lib().orc_set_polygons(len(ps), nv.ctypes.data_as(ip32), _dp(xy))


def box_info(rect=None, polys=None):
    """per box: area, centre (cx, cy), bounding box and border count"""
    rect = _f(poly_rect(polys) if polys is not None else rect)
    out = np.zeros((len(rect), 8))
    set_polygons(polys)
    if lib().orc_box_info(len(rect), _dp(rect), _dp(out)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def borders(ibox, rect=None, polys=None, maxb=32):
    """borders of box ibox in the oracle's order: rows p1x p1y p2x p2y npx npy"""
    rect = _f(poly_rect(polys) if polys is not None else rect)
    out = np.zeros((maxb, 6))
    set_polygons(polys)
    n = lib().orc_borders(len(rect), _dp(rect), ibox, maxb, _dp(out))
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out[:n]


def inside(ibox, xy, rect=None, polys=None):
    rect = _f(poly_rect(polys) if polys is not None else rect)
    xy = _f(np.asarray(xy, np.float64).reshape(-1, 2))
    out = np.zeros(len(xy), np.int32)
    set_polygons(polys)
    if lib().orc_inside(len(rect), _dp(rect), ibox, len(xy), _dp(xy), out.ctypes.data_as(ip32)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out.astype(bool)


def random_points(ibox, n, seed, rect=None, polys=None):
    rect = _f(poly_rect(polys) if polys is not None else rect)
    out = np.zeros((n, 2))
    set_polygons(polys)
    if lib().orc_random_points(len(rect), _dp(rect), ibox, n, C.c_uint64(seed), _dp(out)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def set_periodic(nbox, periodic):
    """periodic: {box: (target, tx, ty)} for the next oracle run"""
    if not periodic:
        return
    tgt = np.full(nbox, -1, np.int32)
    tr = np.zeros(2 * nbox)
    for b, (t, tx, ty) in periodic.items():
        tgt[b] = t
        tr[2 * b:2 * b + 2] = (tx, ty)
    lib().orc_set_periodic(nbox, tgt.ctypes.data_as(ip32), _dp(tr))


def rta_run(mat, rect, kind, Teq, thickness, T0, dt, maxtime, nparticles, seed, ballistic=False, sigma=0.25,
            nthreads=0, periodic=None, spectral=None, wscale=1.0, specularity=None, grad=False, mode_sigma=None,
            polys=None):
    """spectral: {"boxes": [...], "nomega": K, "sigma": per-mode array or
    None}; the result then holds "spectral": one dict per location with
    omega, dT, jx, jy [nt][K] and fd [nt][nq]. wscale multiplies every
    scattering rate; specularity (left, bottom, top, right) is the specular
    probability of each side; grad adds "dTdp" [4][nt][nbox] for the sides
    with 0 < p < 1 (zeros elsewhere). mode_sigma: per-mode base broadening of
    the Lambert bounce (None: sigma for every mode)."""
    rect, Teq = _f(poly_rect(polys) if polys is not None else rect), _f(Teq)
    kind = np.ascontiguousarray(kind, np.int32)
    nbox = len(kind)
    set_periodic(nbox, periodic)
    set_polygons(polys)
    nq = int(mat.get("nq", 1))
    if spectral:
        sb = np.ascontiguousarray(spectral["boxes"], np.int32)
        sg = spectral.get("sigma")
        lib().orc_rta_spectral(len(sb), sb.ctypes.data_as(ip32), int(spectral["nomega"]), len(mat["omega"]),
                               None if sg is None else _dp(_f(sg)))
    ntime = int(np.floor(maxtime / dt + 1e-6)) + 1
    nt = ntime - 1
    T = np.zeros((nt, nbox))
    jx = np.zeros((nt, nbox))
    jy = np.zeros((nt, nbox))
    raw = np.zeros((nbox, nt))
    cnt = (C.c_uint64 * 3)()
    eeff = C.c_double()
    om, vx, vy, w = (_f(mat[k]) for k in ("omega", "vx", "vy", "w"))
    sp = _f(specularity if specularity is not None else (0.0, 0.0, 0.0, 0.0))
    lib().orc_rta_ext(C.c_double(wscale), _dp(sp), 15 if grad else 0)
    ms = None if mode_sigma is None else _f(mode_sigma)
    lib().orc_rta_mode_sigma(len(mat["omega"]), None if ms is None else _dp(ms))
    rc = lib().orc_rta_run(len(om), _dp(om), _dp(vx), _dp(vy), _dp(w), C.c_double(mat["V"]),
                           C.c_double(mat.get("lattz", 1.0)), nq, nbox, _dp(rect), kind.ctypes.data_as(ip32),
                           _dp(Teq), C.c_double(thickness), C.c_double(T0), C.c_double(dt), C.c_double(maxtime),
                           u64(nparticles), u64(seed), int(ballistic), C.c_double(sigma), nthreads, _dp(T),
                           _dp(jx), _dp(jy), cnt, _dp(raw), C.byref(eeff))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = {"T": T, "jx": jx, "jy": jy, "events": cnt[0], "crossings": cnt[1], "nudges": cnt[2], "raw_gz": raw,
           "Eeff": eeff.value}
    if grad:
        g = np.zeros((4, nt, nbox))
        if lib().orc_rta_grad_get(_dp(g)) != 0:
            raise RuntimeError("oracle returned no dT/dp")
        out["dTdp"] = g
    if spectral:
        K = int(spectral["nomega"])
        out["spectral"] = []
        for k in range(len(spectral["boxes"])):
            d = {"omega": np.zeros(K), "dT": np.zeros((nt, K)), "jx": np.zeros((nt, K)), "jy": np.zeros((nt, K)),
                 "fd": np.zeros((nt, nq))}
            lib().orc_rta_spectral_get(k, *(_dp(d[x]) for x in ("omega", "dT", "jx", "jy", "fd")))
            out["spectral"].append(d)
    return out


def gamma_pdf(k, theta, x):
    return lib().orc_gamma_pdf(C.c_double(k), C.c_double(theta), C.c_double(x))


def rta_steady_run(mat, rect, kind, Teq, thickness, T0, gradient, nparticles, seed, sigma=0.25, eps_energy=-42.0,
                   eps_flux=-42.0, max_generations=0, nthreads=0, periodic=None, polys=None):
    """steady branch: per box T, jx, jy and the number of generations"""
    rect, Teq = _f(poly_rect(polys) if polys is not None else rect), _f(Teq)
    kind = np.ascontiguousarray(kind, np.int32)
    nbox = len(kind)
    set_periodic(nbox, periodic)
    set_polygons(polys)
    T, jx, jy = np.zeros(nbox), np.zeros(nbox), np.zeros(nbox)
    gens = C.c_int32()
    eeff = C.c_double()
    om, vx, vy, w = (_f(mat[k]) for k in ("omega", "vx", "vy", "w"))
    rc = lib().orc_rta_steady_run(len(om), _dp(om), _dp(vx), _dp(vy), _dp(w), C.c_double(mat["V"]),
                                  C.c_double(mat.get("lattz", 1.0)), 1, nbox, _dp(rect), kind.ctypes.data_as(ip32),
                                  _dp(Teq), C.c_double(thickness), C.c_double(T0), C.c_double(gradient[0]),
                                  C.c_double(gradient[1]), C.c_double(eps_energy), C.c_double(eps_flux),
                                  int(max_generations), u64(nparticles), u64(seed), C.c_double(sigma), nthreads,
                                  _dp(T), _dp(jx), _dp(jy), C.byref(gens), C.byref(eeff))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return {"T": T, "jx": jx, "jy": jy, "generations": gens.value, "Eeff": eeff.value}


def bert_run(mat, rect, kind, Teq, Tinit, thickness, dt, maxtime, P=None, ngamma=0, rta=False, Ed=-1.0,
             Efactor=0, Tmax=0.0, Tmin=0.0, seed=1, max_steps=0, nthreads=0, periodic=None, Ps=None, Ts=None,
             prop_rule=0, gradient=None, pcancel_each=4, init_dist=None, polys=None):
    """Ps/Ts: propagators at several temperatures (replace P); gradient:
    (gx, gy) in every regular box instead of calculate_gradientT; init_dist:
    {box: signed distribution over the modes after the ngamma offset}"""
    rect, Teq, Tinit = _f(poly_rect(polys) if polys is not None else rect), _f(Teq), _f(Tinit)
    kind = np.ascontiguousarray(kind, np.int32)
    nbox = len(kind)
    if Ps is not None:
        ncol = Ps[0].shape[0]
        Pa = np.concatenate([np.asfortranarray(x, dtype=np.float64).ravel(order="F") for x in Ps])
        P = Ps[0]
    else:
        ncol, Pa = 1, np.zeros(1)
    gx, gy = gradient if gradient is not None else (0.0, 0.0)
    lib().orc_bert_options(ncol, 0 if Ps is None else len(Ps), _dp(_f(np.asarray(Ts if Ts is not None else [0.0],
                                                                              np.float64))),
                           _dp(Pa), prop_rule, int(gradient is not None), C.c_double(gx), C.c_double(gy),
                           int(pcancel_each))
    set_periodic(nbox, periodic)
    set_polygons(polys)
    n = len(mat["omega"])
    if init_dist:
        ib = np.ascontiguousarray(sorted(init_dist), np.int32)
        lens = {len(init_dist[b]) for b in ib}
        if len(lens) != 1:
            raise ValueError("given distributions of different lengths")
        fd = np.ascontiguousarray(np.concatenate([np.asarray(init_dist[b], np.float64) for b in ib]))
        lib().orc_bert_init_dist(len(ib), ib.ctypes.data_as(ip32), lens.pop(), _dp(fd))
    else:
        lib().orc_bert_init_dist(0, None, 0, None)
    nst = 0
    t = 0.0
    while t <= maxtime:
        nst += 1
        t += dt
    if max_steps > 0:
        nst = min(nst, max_steps)
    props = np.zeros((nst + 1, nbox, 3))
    npart = np.zeros(nst + 1, np.int64)
    pending = np.zeros(nst + 1, np.int64)
    H = np.zeros((nbox, n), np.int32)
    eeff = C.c_double()
    ns = C.c_int32()
    Pm = np.asfortranarray(P, dtype=np.float64) if P is not None else np.zeros((1, 1))
    om, vx, vy, w = (_f(mat[k]) for k in ("omega", "vx", "vy", "w"))
    rc = lib().orc_bert_run(n, _dp(om), _dp(vx), _dp(vy), _dp(w), C.c_double(mat["V"]),
                            C.c_double(mat.get("lattz", 1.0)), 1, nbox, _dp(rect), kind.ctypes.data_as(ip32),
                            _dp(Teq), _dp(Tinit), C.c_double(thickness), C.c_double(dt), C.c_double(maxtime),
                            ngamma, Pm.ctypes.data_as(dp), int(rta), C.c_double(Ed), u64(Efactor),
                            C.c_double(Tmax), C.c_double(Tmin), u64(seed), nthreads, nst, _dp(props),
                            npart.ctypes.data_as(C.POINTER(C.c_int64)), H.ctypes.data_as(ip32), C.byref(eeff),
                            C.byref(ns), pending.ctypes.data_as(C.POINTER(C.c_int64)))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    k = ns.value
    return {"props": props[:k + 1], "npart": npart[:k + 1], "pending": pending[:k + 1], "H_last": H,
            "Eeff": eeff.value, "nsteps": k}


def expm(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    E = np.zeros((n, n), order="F")
    lib().orc_expm(n, A.ctypes.data_as(dp), E.ctypes.data_as(dp))
    return E


def arnoldi(A, b, n, t, mode):
    A = np.asfortranarray(A, dtype=np.float64)
    N = A.shape[0]
    b = _f(b)
    V = np.zeros((N, n), order="F")
    H = np.zeros((n, n), order="F")
    lib().orc_arnoldi(N, A.ctypes.data_as(dp), _dp(b), n, C.c_double(t), mode, V.ctypes.data_as(dp),
                      H.ctypes.data_as(dp))
    return V, H


def build_P(A, t, m, mode, nthreads=0):
    A = np.asfortranarray(A, dtype=np.float64)
    N = A.shape[0]
    P = np.zeros((N, N), order="F")
    lib().orc_build_P(N, A.ctypes.data_as(dp), C.c_double(t), m, mode, nthreads, P.ctypes.data_as(dp))
    return P


def build_P_cols(A, t, m, mode, c0, c1, nthreads=0):
    """columns [c0, c1) of the per-column Arnoldi build"""
    A = np.asfortranarray(A, dtype=np.float64)
    N = A.shape[0]
    P = np.zeros((N, c1 - c0), order="F")
    lib().orc_build_P_cols(N, A.ctypes.data_as(dp), C.c_double(t), m, mode, c0, c1, nthreads,
                           P.ctypes.data_as(dp))
    return P


def krylov_size(A, t, eps=1e-8):
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().orc_krylov_size(A.shape[0], A.ctypes.data_as(dp), C.c_double(t), C.c_double(eps))


def error_bound(A, n, t):
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().orc_error_bound(A.shape[0], A.ctypes.data_as(dp), n, C.c_double(t))


def synthetic_B(Cm, w, density, seed, iters=200):
    Cm, w = _f(Cm), _f(w)
    n = len(Cm)
    B = np.zeros((n, n), order="F")
    lib().orc_synthetic_B(n, _dp(Cm), _dp(w), C.c_double(density), u64(seed), iters, B.ctypes.data_as(dp))
    return B


def project_columns(B):
    B = np.asfortranarray(B, dtype=np.float64).copy(order="F")
    lib().orc_project_columns(B.shape[0], B.ctypes.data_as(dp))
    return B


def lambert_sample(omega, vx, vy, sigma, npx, npy, mode, seed, ns, mode_sigma=None):
    """ns diffuse bounces of `mode` at a side with outward normal (npx, npy);
    mode_sigma: per-mode base broadening (None: sigma for every mode)."""
    om, ux, uy = _f(omega), _f(vx), _f(vy)
    out = np.zeros(ns, np.int32)
    ms = None if mode_sigma is None else _f(mode_sigma)
    if lib().orc_lambert_sample(len(om), _dp(om), _dp(ux), _dp(uy), C.c_double(sigma), C.c_double(npx),
                                C.c_double(npy), int(mode), u64(seed), int(ns),
                                out.ctypes.data_as(ip32), None if ms is None else _dp(ms)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


EVENT_KINDS = {0: "inside", 1: "bounce", 2: "crossing", 3: "nudge", 4: "absorbed", 5: "translated"}


def rta_event(mat, rect, kind, ibox, x, y, mode, dt, seed, ns, sigma=0.25):
    """ns repeats (streams 0..ns-1) of one ballistic evolparticle event."""
    om, ux, uy = _f(mat["omega"]), _f(mat["vx"]), _f(mat["vy"])
    rc = _f(rect).reshape(-1, 4)
    kd = np.ascontiguousarray(kind, np.int32)
    box = np.zeros(ns, np.int32)
    xy = np.zeros((ns, 2))
    md = np.zeros(ns, np.int32)
    dtl = np.zeros(ns)
    kk = np.zeros(ns, np.int32)
    if lib().orc_rta_event(len(om), _dp(om), _dp(ux), _dp(uy), C.c_double(sigma), len(rc), _dp(rc),
                           kd.ctypes.data_as(ip32), int(ibox), C.c_double(x), C.c_double(y), int(mode),
                           C.c_double(dt), u64(seed), int(ns), box.ctypes.data_as(ip32), _dp(xy),
                           md.ctypes.data_as(ip32), _dp(dtl), kk.ctypes.data_as(ip32)) != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return {"box": box, "xy": xy, "mode": md, "dt": dtl, "kind": kk}
