"""Synthetic inputs of the benchmark configurations (host-side, no GPU).

A synthetic material stands in for the ab-initio HDF5 database: one q-point
with n flat modes (mode index = band index), omega uniform in f = 0.1-15 THz,
|v| uniform 0.5-8 nm/ps at a uniform in-plane angle, tau log-uniform
1 ps - 1 ns. The unit-cell volume V is chosen so that the volumetric heat
capacity at the reference temperature is 1.6e6 J/(m^3 K); lattvec(2,2)
equals the thickness so the reference's thickness correction is 1.
"""
import numpy as np

KB = 1.3806488e-23
HBAR = 1.054571726e-34
CV_TARGET = 1.6e-21  # J/(K nm^3) = 1.6e6 J/(m^3 K)


def be_kernel(omega, T):
    x = 0.5e12 * HBAR * np.asarray(omega) / (KB * T)
    with np.errstate(over="ignore", invalid="ignore"):
        s = x / np.sinh(x)
    s = np.where(x < 1e-320, 1.0, s)
    return np.where(x > 300.0, 0.0, s * s)


def material(n, seed=1, T=300.0, thickness=1.0, mirror=False):
    from . import synthetic_modes
    m = synthetic_modes(n, seed, mirror=mirror)
    ker = be_kernel(m["omega"], T)
    m["V"] = KB * ker.sum() / CV_TARGET
    m["lattz"] = thickness
    m["C"] = KB * ker  # per-mode heat capacity (per unit cell)
    m["n"] = n
    return m


def ribbon(length, width, nbins, Teq=300.0, reservoirs=None, Tinit_first=None):
    """x-binned rectangle [0, length] x [0, width]; reservoirs=(Tleft, Tright)
    adds two reservoir boxes of one bin width at the ends (ids nbins, nbins+1).
    Returns rect (nbox, 4), kind, Teq, Tinit."""
    dx = length / nbins
    rect, kind, teq, tin = [], [], [], []
    for k in range(nbins):
        rect.append([k * dx, (k + 1) * dx, 0.0, width])
        kind.append(0)
        teq.append(Teq)
        tin.append(Teq)
    if Tinit_first is not None:
        tin[0] = Tinit_first
    if reservoirs is not None:
        rect.append([-dx, 0.0, 0.0, width])
        kind.append(1)
        teq.append(reservoirs[0])
        tin.append(reservoirs[0])
        rect.append([length, length + dx, 0.0, width])
        kind.append(1)
        teq.append(reservoirs[1])
        tin.append(reservoirs[1])
    return (np.array(rect, float), np.array(kind, np.int32), np.array(teq, float),
            np.array(tin, float))


def config0(nparticles=1_000_000, nmodes=3000, nbins=100, maxtime=5000.0, dt=25.0, seed=1):
    """RTAMC2D steady state: 1 um x 200 nm, isothermal x-walls at +-0.5 K,
    diffuse top/bottom, 100 x-bins; dt = maxtime/200."""
    mat = material(nmodes, seed=1)
    rect, kind, teq, _ = ribbon(1000.0, 200.0, nbins, 300.0, reservoirs=(300.5, 299.5))
    return dict(material=mat, rect=rect, kind=kind, Teq=teq, thickness=1.0, T0=300.0, dt=dt,
                maxtime=maxtime, nparticles=nparticles, seed=seed, sigma_bounce=0.25)


def config1_inputs(nmodes=2000, seed=1, density=0.05):
    """beRTAMC2D material and collision operator B = -Omega (dense,
    column-major) for the 2000-mode synthetic table."""
    from . import synthetic_B
    mat = material(nmodes, seed=seed)
    B = synthetic_B(mat["C"], mat["w"], density=density, seed=seed)
    return mat, B


def config1(nsteps=200, dt=5.0, nparticles=1_000_000, nbins=50, seed=1, replicas=1, gap=100.0):
    """beRTAMC2D transient: 500 nm x 100 nm ribbon, 50 x-bins, 1 K pulse in
    bin 0, all sides adiabatic-diffuse, Eeff set for ~nparticles initial
    particles per ribbon, dt = 5 ps, nsteps steps. replicas > 1 places that
    many copies of the ribbon side by side in y, gap nm apart: they share no
    side, so one run simulates independent replicas (boxes r*nbins ...
    (r+1)*nbins - 1 belong to replica r)."""
    parts = [ribbon(500.0, 100.0, nbins, 300.0, Tinit_first=301.0) for _ in range(replicas)]
    for r, pr in enumerate(parts):
        pr[0][:, 2:] += r * (100.0 + gap)
    rect = np.concatenate([p[0] for p in parts])
    kind = np.concatenate([p[1] for p in parts])
    teq = np.concatenate([p[2] for p in parts])
    tin = np.concatenate([p[3] for p in parts])
    return dict(rect=rect, kind=kind, Teq=teq, Tinit=tin, thickness=1.0, dt=dt,
                maxtime=dt * (nsteps - 1), max_steps=nsteps, Efactor=nparticles, Tmax=301.0,
                Tmin=300.0, seed=seed)


def config3(nruns=64, nmodes=3000, nbins=100, length=1000.0, maxtime=5000.0, dt=25.0, seed=1):
    """Batched sweep of independent RTAMC2D runs: widths log-spaced
    50 nm - 2 um, temperatures 100 - 500 K (walls at T +- 0.5 K), Cv and the
    BE_derivative sampler at each run's T, scattering rates scaled as
    tau(T) = tau(300 K) * 300 / T, particle counts log-uniform 1e5 - 1e7
    (seeded), the 3000-mode table of config 0 shared by all runs."""
    mat = material(nmodes, seed=1)
    widths = np.logspace(np.log10(50.0), np.log10(2000.0), nruns)
    temps = np.linspace(100.0, 500.0, nruns)
    rng = np.random.default_rng(seed)
    counts = np.round(10.0 ** rng.uniform(5.0, 7.0, nruns)).astype(np.int64)
    runs = []
    for r in range(nruns):
        rect, kind, teq, _ = ribbon(length, widths[r], nbins, temps[r],
                                    reservoirs=(temps[r] + 0.5, temps[r] - 0.5))
        runs.append(dict(rect=rect, kind=kind, Teq=teq, T0=float(temps[r]), nparticles=int(counts[r]),
                         seed=seed * 1000 + r, wscale=float(temps[r]) / 300.0, width=float(widths[r])))
    return dict(material=mat, runs=runs, thickness=1.0, dt=dt, maxtime=maxtime, length=length,
                sigma_bounce=0.25)


def config4(nparticles=10_000_000, nmodes=5000, nbins=128, ntime=64, maxtime=5000.0, p_top=0.3,
            p_bottom=0.7, seed=1):
    """Differentiable transient RTAMC2D: 1 um x 100 nm ribbon in 128 x-bins,
    64 time bins, walls at 300 +- 0.5 K, specular probability p_top on the
    top edge and p_bottom on the bottom edge (diffuse otherwise), 5000 modes
    built as 2500 mirror pairs in v_y."""
    mat = material(nmodes, seed=seed, mirror=True)
    rect, kind, teq, _ = ribbon(1000.0, 100.0, nbins, 300.0, reservoirs=(300.5, 299.5))
    # side order: left, bottom, top, right
    spec = (0.0, p_bottom, p_top, 0.0)
    return dict(material=mat, rect=rect, kind=kind, Teq=teq, T0=300.0, thickness=1.0, dt=maxtime / ntime,
                maxtime=maxtime, nparticles=nparticles, seed=seed, specularity=spec, sigma_bounce=0.25)
