"""ctypes binding of libbte_amd.so (the C ABI declared in include/bte_abi.h).

The library is the product path; there is no CPU fallback. Loading fails
loudly when the shared object has not been built."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libbte_amd.so")

BTE_OK = 0
ERRORS = {-1: "EINVAL", -2: "EGEOM", -3: "EHIP", -4: "ENOMEM", -5: "ECAPACITY", -6: "ENODEV"}

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
ARNOLDI = {"reference": 0, "textbook": 1}


class BteError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class RtaCfg(C.Structure):
    _fields_ = [("thickness", C.c_double), ("T0", C.c_double), ("dt", C.c_double),
                ("maxtime", C.c_double), ("nparticles", C.c_uint64), ("seed", C.c_uint64),
                ("ballistic", C.c_int), ("sigma_bounce", C.c_double)]


class RtaReport(C.Structure):
    _fields_ = [("events", C.c_uint64), ("crossings", C.c_uint64), ("Eeff", C.c_double),
                ("kernel_ms", C.c_double), ("ntime", C.c_int)]


class RtaSteadyCfg(C.Structure):
    _fields_ = [("thickness", C.c_double), ("T0", C.c_double), ("gradient", C.c_double * 2),
                ("nparticles", C.c_uint64), ("seed", C.c_uint64), ("sigma_bounce", C.c_double),
                ("eps_energy", C.c_double), ("eps_flux", C.c_double), ("max_generations", C.c_int)]


class RtaSteadyReport(C.Structure):
    _fields_ = [("generations", C.c_int), ("converged", C.c_int), ("flights", C.c_uint64), ("Eeff", C.c_double),
                ("kernel_ms", C.c_double)]


class RtaRunDesc(C.Structure):
    _fields_ = [("geom", C.c_void_p), ("T0", C.c_double), ("nparticles", C.c_uint64), ("seed", C.c_uint64),
                ("ballistic", C.c_int), ("wscale", C.c_double), ("specularity", C.c_double * 4),
                ("grad", C.c_int)]


class RtaBatchReport(C.Structure):
    _fields_ = [("events", C.c_uint64), ("crossings", C.c_uint64), ("kernel_ms", C.c_double),
                ("ntime", C.c_int), ("block", C.c_int), ("blocks", C.c_int)]


class BertCfg(C.Structure):
    _fields_ = [("thickness", C.c_double), ("dt", C.c_double), ("maxtime", C.c_double),
                ("max_steps", C.c_int), ("Ed", C.c_double), ("Efactor", C.c_uint64),
                ("Tmax", C.c_double), ("Tmin", C.c_double), ("rta", C.c_int),
                ("pcancel_each", C.c_int), ("seed", C.c_uint64), ("capacity", C.c_uint64),
                ("fixed_gradient", C.c_int), ("gradient", C.c_double * 2), ("prop_interp", C.c_int),
                ("ninit", C.c_int), ("init_len", C.c_int), ("init_boxes", C.POINTER(C.c_int32)),
                ("init_dist", C.POINTER(C.c_double)), ("step_shape", C.c_int)]


class BertReport(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("Eeff", C.c_double), ("particle_steps", C.c_uint64),
                ("step_kernel_particle_steps", C.c_uint64), ("max_particles", C.c_int64),
                ("loop_ms", C.c_double), ("step_kernel_ms", C.c_double), ("step_kernel_launches", C.c_int),
                ("step_kernel_timed", C.c_int), ("step_threads", C.c_int)]


class KrylovReport(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("gemm_ms", C.c_double), ("gemm_tflops", C.c_double),
                ("max_col_sum_err", C.c_double), ("breakdowns", C.c_int)]


FRAME_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                       C.POINTER(C.c_int32))

dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
up32 = C.POINTER(C.c_uint32)
ip64 = C.POINTER(C.c_int64)
vp = C.c_void_p
pvp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every entry point of include/bte_abi.h
SIGNATURES = {
    "bte_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "bte_version": (C.c_int, []),
    "bte_ctx_create": (C.c_int, [C.c_int, pvp]),
    "bte_ctx_destroy": (C.c_int, [vp]),
    "bte_sync": (C.c_int, [vp]),
    "bte_synthetic_modes": (C.c_int, [C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp]),
    "bte_synthetic_B": (C.c_int, [C.c_int, dp, dp, C.c_double, C.c_uint64, C.c_int, dp]),
    "bte_alias_build": (C.c_int, [C.c_int, dp, up32, up32]),
    "bte_modes_create": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp, dp, C.c_double, C.c_double, pvp]),
    "bte_modes_set_sigma": (C.c_int, [vp, dp]),
    "bte_modes_destroy": (C.c_int, [vp]),
    "bte_geom_create_rects": (C.c_int, [vp, C.c_int, dp, ip32, dp, dp, pvp]),
    "bte_geom_create_polys": (C.c_int, [vp, C.c_int, ip32, dp, ip32, dp, dp, pvp]),
    "bte_geom_destroy": (C.c_int, [vp]),
    "bte_geom_set_periodic": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double]),
    "bte_geom_contacts": (C.c_int, [vp, C.c_int, ip32, ip32, dp]),
    "bte_geom_gradient_T": (C.c_int, [vp, dp]),
    "bte_rta_create": (C.c_int, [vp, vp, vp, C.POINTER(RtaCfg), pvp]),
    "bte_rta_execute": (C.c_int, [vp, dp, dp, dp, C.POINTER(RtaReport)]),
    "bte_rta_execute_noalloc": (C.c_int, [vp, vp, vp, vp, vp]),
    "bte_rta_destroy": (C.c_int, [vp]),
    "bte_rta_spectral": (C.c_int, [vp, C.c_int, ip32, C.c_int, dp]),
    "bte_rta_spectral_get": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp]),
    "bte_spectra": (C.c_int, [vp, vp, dp, C.c_double, C.c_int, dp, C.c_int, dp, dp]),
    "bte_expm": (C.c_int, [vp, C.c_int, dp, C.c_double, dp, dp]),
    "bte_rta_run": (C.c_int, [vp, vp, vp, C.POINTER(RtaCfg), dp, dp, dp, C.POINTER(RtaReport)]),
    "bte_rta_steady_run": (C.c_int, [vp, vp, vp, C.POINTER(RtaSteadyCfg), dp, dp, dp, C.POINTER(RtaSteadyReport)]),
    "bte_rta_batch_create": (C.c_int, [vp, vp, C.c_int, C.POINTER(RtaRunDesc), C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, pvp]),
    "bte_rta_batch_set_tally": (C.c_int, [vp, C.c_int]),
    "bte_rta_batch_tune": (C.c_int, [vp, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bte_rta_batch_execute": (C.c_int, [vp, dp, dp, dp, dp, dp, C.POINTER(RtaBatchReport)]),
    "bte_rta_batch_destroy": (C.c_int, [vp]),
    "bte_prop_create": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_double, pvp]),
    "bte_prop_create_multi": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, C.c_int, pvp]),
    "bte_prop_pick_pmf": (C.c_int, [C.c_int, dp, dp]),
    "bte_prop_destroy": (C.c_int, [vp]),
    "bte_bert_create": (C.c_int, [vp, vp, vp, vp, C.POINTER(BertCfg), pvp]),
    "bte_bert_execute": (C.c_int, [vp, dp, ip64, ip64, FRAME_CB, vp, C.POINTER(BertReport)]),
    "bte_bert_destroy": (C.c_int, [vp]),
    "bte_krylov_build": (C.c_int, [vp, C.c_int, dp, C.c_double, C.c_int, C.c_int, C.c_int, dp,
                                   C.POINTER(KrylovReport)]),
    "bte_krylov_size": (C.c_int, [C.c_int, dp, C.c_double, C.c_double, C.POINTER(C.c_int)]),
}

_lib = None


def lib():
    """Load libbte_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().bte_last_error(buf, len(buf))
    return buf.value.decode()


def check(rc):
    if rc != BTE_OK:
        raise BteError(rc, last_error())
    return rc
