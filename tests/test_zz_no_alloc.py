"""bte_rta_execute_noalloc, the allocation-free variant of bte_rta_execute
(its tally copies, counters and per-workgroup blocks in the library's static
device scratch array), against bte_rta_execute on the same handles: two calls
in a row, the second on a smaller run, each with the same histories (equal
event and crossing counts) and T, jx, jy equal up to the order of the FP64
sums. A run whose tallies exceed the scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself links (torch's
    bundled runtime cannot share the device with it in one process)."""
    hip = None

    def __init__(self, n, dtype):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = C.CDLL("libamdhip64.so.7")
        self.n, self.dtype = n, np.dtype(dtype)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(n * self.dtype.itemsize)) == 0
        zero = np.zeros(n, self.dtype)
        assert self.hip.hipMemcpy(self.ptr, zero.ctypes.data_as(C.c_void_p), C.c_size_t(zero.nbytes), 1) == 0

    def get(self):
        out = np.zeros(self.n, self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_noalloc_matches_execute(pkg, ctx):
    W = pkg.workloads
    mat = W.material(400, seed=6)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    cases = [(W.ribbon(600.0, 100.0, 60, 300.0, reservoirs=(300.5, 299.5)), 200000, 5.0, 1000.0),
             (W.ribbon(100.0, 40.0, 8, 300.0, reservoirs=(301.0, 299.0)), 30000, 10.0, 300.0)]
    for (rect, kind, teq, _), N, dt, tmax in cases:
        geom = pkg.Geometry(ctx, rect, kind, teq)
        h = pkg.RtaRun(ctx, modes, geom, thickness=1.0, T0=300.0, dt=dt, maxtime=tmax, nparticles=N, seed=9,
                       sigma_bounce=2.0)
        ref = h.execute()
        nt, nb = ref["T"].shape
        T, jx, jy = (DeviceBuffer(nt * nb, np.float64) for _ in range(3))
        cnt = DeviceBuffer(6, np.uint64)
        print(f"noalloc: {nb} boxes x {nt} bins, {N} particles", flush=True)
        h.execute_noalloc(T.ptr.value, jx.ptr.value, jy.ptr.value, cnt.ptr.value)
        ctx.sync()
        c = cnt.get()
        assert (c[0], c[1]) == (ref["events"], ref["crossings"]) and c[3:].sum() == 0, (c, ref["events"])
        for key, t, off in (("T", T, 300.0), ("jx", jx, 0.0), ("jy", jy, 0.0)):
            got = t.get().reshape(nt, nb)
            scale = np.abs(ref[key] - off).max()
            assert scale > 0
            np.testing.assert_allclose(got - off, ref[key] - off, rtol=0, atol=1e-10 * scale)
        for b in (T, jx, jy, cnt):
            b.free()
        h.close()
        geom.close()
    # 500 boxes x 2000 time bins: 8 copies x 3 tallies x 1e6 cells x 8 B > 64 MB
    rect, kind, teq, _ = W.ribbon(5000.0, 20.0, 498, 300.0, reservoirs=(300.5, 299.5))
    geom = pkg.Geometry(ctx, rect, kind, teq)
    h = pkg.RtaRun(ctx, modes, geom, thickness=1.0, T0=300.0, dt=1.0, maxtime=2000.0, nparticles=1000, seed=1,
                   sigma_bounce=2.0)
    T = DeviceBuffer(2000 * 500, np.float64)
    cnt = DeviceBuffer(6, np.uint64)
    with pytest.raises(pkg.BteError, match="scratch"):
        h.execute_noalloc(T.ptr.value, T.ptr.value, T.ptr.value, cnt.ptr.value)
    T.free()
    cnt.free()
    h.close()
