"""The C++ front-ends on the GPU: PropagatorBuilder against the oracle's
Krylov build, beRTAMC2D stdout/frames against each other and against the
library run with the same seed, RTAMC2D CSVs against the library run."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu

GEOM = """<Geometry>
  <number_of_boxes Ngeom="{n}"/>
{boxes}
</Geometry>"""
BOX = """  <Box>
    <BoxID id="{i}"/>
    <MaterialID name="m"/>
    <Vertices dim="2" npoints="4"> {x0} {y0} {x0} {y1} {x1} {y1} {x1} {y0} </Vertices>
    <initCnd Teq="{teq}" Tinit="{tin}"/>{extra}
  </Box>"""


def write_ribbon(path, rect, kind, teq, tin):
    boxes = [BOX.format(i=i, x0=r[0], x1=r[1], y0=r[2], y1=r[3], teq=teq[i], tin=tin[i],
                        extra="\n    <Reservoir/>" if kind[i] == 1 else "") for i, r in enumerate(rect)]
    path.write_text(GEOM.format(n=len(rect), boxes="\n".join(boxes)))


def run(cmd, cwd):
    print("front-end:", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    return r


def eigen_bin(path):
    b = path.read_bytes()
    n = int(np.frombuffer(b[:8], np.int64)[0])
    return np.frombuffer(b[8:], np.float64).reshape(n, n, order="F")


def test_propagator_builder_matches_oracle(pkg, tmp_path):
    run([pkg.app_path("PropagatorBuilder"), "synthetic:96:2", "300", "5", "1", "20", "fp64", "reference"], tmp_path)
    B = eigen_bin(tmp_path / "synthetic_96_2_300K_5ps.B.eigen.bin")
    P = eigen_bin(tmp_path / "synthetic_96_2_300K_5ps.P.eigen.bin")
    mat, B2 = pkg.workloads.config1_inputs(nmodes=96, seed=2)
    # same operator up to libm-vs-numpy rounding of the Bose-Einstein kernel
    assert np.allclose(B, B2, rtol=1e-12, atol=1e-15 * np.abs(B2).max())
    Po = orc.build_P(B, 5.0, 20, 0)
    assert np.abs(P - Po).max() < 1e-10
    assert np.abs(P.sum(0) - 1).max() < 1e-12


def test_propagator_builder_pade(pkg, tmp_path):
    """The "pade" method writes exp(dt B) to rounding: the oracle's Pade
    exponential and, within the Krylov error, the oracle's Krylov build"""
    run([pkg.app_path("PropagatorBuilder"), "synthetic:96:2", "300", "5", "1", "0", "pade"], tmp_path)
    B = eigen_bin(tmp_path / "synthetic_96_2_300K_5ps.B.eigen.bin")
    P = eigen_bin(tmp_path / "synthetic_96_2_300K_5ps.P.eigen.bin")
    assert np.abs(P - orc.expm(5.0 * B)).max() < 1e-13
    assert np.abs(P - orc.build_P(B, 5.0, 20, 0)).max() < 1e-10
    assert np.abs(P.sum(0) - 1).max() < 1e-12


def test_bertamc2d_outputs(pkg, ctx, tmp_path):
    msgpack = pytest.importorskip("msgpack")
    W = pkg.workloads
    run([pkg.app_path("PropagatorBuilder"), "synthetic:96:2", "300", "5", "1", "20"], tmp_path)
    rect, kind, teq, tin = W.ribbon(60.0, 30.0, 6, 300.0, Tinit_first=301.0)
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, tin)
    (tmp_path / "in.xml").write_text(
        '<beRTAMC2D>\n<geometry file="g.xml"/>\n<time dt="5" maxtime="45"/>\n'
        '<material name="m" modetable="synthetic" nmodes="96" seed="2" thickness="1" T0="300">\n'
        '  <propagator T="300" file="synthetic_96_2_300K_5ps.P.eigen.bin"/>\n</material>\n'
        '<Eeff particles="5000" Tmax="301" Tmin="300"/>\n<seed value="7"/>\n</beRTAMC2D>\n')
    r = run([pkg.app_path("beRTAMC2D"), "in.xml", "4"], tmp_path)
    lines = [ln.split("\t") for ln in r.stdout.splitlines() if ln and not ln.startswith("#")]
    nsteps = 10
    assert len(lines) == nsteps + 1
    rows = np.array([[float(x) for x in ln if x != ""] for ln in lines])
    assert rows.shape == (nsteps + 1, 3 + 3 * 6)
    assert np.allclose(rows[:, 1], 5.0 * np.arange(nsteps + 1))
    # frames: one per step, time (k+1) dt, Teq per box, histogram per box
    raw = (tmp_path / "properties.msgpack.bin").read_bytes()
    frames, pos = [], 0
    while True:
        k = raw.index(b"#", pos)
        size = int(raw[pos:k])
        if size == 0:
            break
        frames.append(msgpack.unpackb(raw[k + 1:k + 1 + size]))
        pos = k + 1 + size
    assert len(frames) == nsteps
    mat = W.material(96, seed=2)
    vx = mat["vx"]
    vol = (rect[:, 1] - rect[:, 0]) * (rect[:, 3] - rect[:, 2]) * 1.0
    for s, f in enumerate(frames):
        assert f[0] == pytest.approx(5.0 * (s + 1)) and f[2] == list(teq)
        Eeff = f[1]
        for b in range(6):
            n2, pairs = f[3 + 2 * b], f[4 + 2 * b]
            h = np.zeros(96)
            if n2:
                h[np.array(pairs[0::2])] = pairs[1::2]
            ed = Eeff / vol[b] * 1e27 * h.sum()
            jx = Eeff / vol[b] * 1e30 * (h * vx).sum()
            assert rows[s + 1, 3 + 3 * b] == pytest.approx(ed, rel=2e-5, abs=1e-300)
            assert rows[s + 1, 4 + 3 * b] == pytest.approx(jx, rel=2e-5, abs=1e-6 * abs(jx) + 1e-300)
    # the library with the same inputs and seed gives the same particle counts
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq, tin)
    P = eigen_bin(tmp_path / "synthetic_96_2_300K_5ps.P.eigen.bin")
    prop = pkg.Propagator(ctx, P)
    o = pkg.bertamc2d(ctx, modes, geom, prop, thickness=1.0, dt=5.0, maxtime=45.0, Efactor=5000, Tmax=301.0,
                      Tmin=300.0, seed=7)
    assert np.array_equal(o["npart"], rows[:, 2].astype(np.int64))
    assert np.allclose(o["props"][:, :, 0], rows[:, 3::3], rtol=2e-5, atol=0)


def test_bertamc2d_temperature_profile(pkg, ctx, tmp_path):
    """Box temperatures 300-305 K: two <propagator> children (300 K and
    306 K, interpolated between), gradient sources from the temperatures or
    from <gradient>; the executable and the library agree on the same seed."""
    W = pkg.workloads
    for T in (300, 306):
        run([pkg.app_path("PropagatorBuilder"), "synthetic:96:2", str(T), "5", "1", "20"], tmp_path)
    rect, kind, _, _ = W.ribbon(60.0, 30.0, 6, 300.0)
    teq = 300.0 + np.arange(6, dtype=float)
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, teq)
    base = ('<beRTAMC2D>\n<geometry file="g.xml"/>\n<time dt="5" maxtime="45"/>\n'
            '<material name="m" modetable="synthetic" nmodes="96" seed="2" thickness="1" T0="300">\n'
            '  <propagator T="300" file="synthetic_96_2_300K_5ps.P.eigen.bin"/>\n'
            '  <propagator T="306" file="synthetic_96_2_306K_5ps.P.eigen.bin"/>\n</material>\n'
            '<Eeff particles="20000" Tmax="305" Tmin="300"/>\n<seed value="3"/>\n{extra}</beRTAMC2D>\n')
    mat = W.material(96, seed=2)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    Ps = [eigen_bin(tmp_path / f"synthetic_96_2_{T}K_5ps.P.eigen.bin") for T in (300, 306)]
    prop = pkg.Propagator(ctx, Ps, T=[300.0, 306.0])
    kw = dict(thickness=1.0, dt=5.0, maxtime=45.0, Efactor=20000, Tmax=305.0, Tmin=300.0, seed=3, prop_interp=1)
    for extra, grad in (('<interpolation rule="bracket"/>\n', None),
                        ('<interpolation rule="bracket"/>\n<gradient x="0.02" y="0.0"/>\n', (0.02, 0.0))):
        (tmp_path / "in.xml").write_text(base.format(extra=extra))
        r = run([pkg.app_path("beRTAMC2D"), "in.xml", "4"], tmp_path)
        lines = [ln.split("\t") for ln in r.stdout.splitlines() if ln and not ln.startswith("#")]
        rows = np.array([[float(x) for x in ln if x != ""] for ln in lines])
        o = pkg.bertamc2d(ctx, modes, geom, prop, gradient=grad, **kw)
        # identical up to the first cancellation of bins holding both signs
        # (which particles survive then depends on their arrival order)
        assert np.array_equal(o["npart"][:2], rows[:2, 2].astype(np.int64))
        assert np.allclose(o["props"][:2, :, 0], rows[:2, 3::3], rtol=2e-5, atol=0)
        assert np.allclose(o["npart"][2:], rows[2:, 2], rtol=0.03)
        assert rows[-1, 2] > 1000  # the gradient sources fill the ribbon
    # Scattering's own pair rule has no pair for 301-305 K with two tables
    (tmp_path / "in.xml").write_text(base.format(extra=""))
    r = subprocess.run([str(pkg.app_path("beRTAMC2D")), "in.xml", "4"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 1 and "Extrapolation" in r.stderr


def test_rtamc2d_csv(pkg, ctx, tmp_path):
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(60.0, 30.0, 6, 300.0, reservoirs=(300.5, 299.5))
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, teq)
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d>\n<geometry file="g.xml"/>\n<time dt="20" maxtime="200"/>\n'
        '<material name="m" modetable="synthetic" nmodes="400" seed="2" thickness="1" T0="300" sigma_bounce="2.0"/>\n'
        '<particles N="20000" seed="5"/>\n</RTA_MC2d>\n')
    # three runs go to the GPU as one batch: the histories of three single
    # runs with seeds 5, 6, 7; only the order of the FP64 tally sums differs
    r = run([pkg.app_path("RTAMC2D"), "in.xml", "8", "3"], tmp_path)
    assert "as one batch" in r.stdout
    mat = W.material(400, seed=2)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    for irun in range(3):
        T = np.loadtxt(tmp_path / f"temperature_300K_run_{irun}.csv", delimiter=",")
        jx = np.loadtxt(tmp_path / f"jx_300K_run_{irun}.csv", delimiter=",")
        assert T.shape == (10, 1 + 8) and jx.shape == T.shape
        assert np.allclose(T[:, 0], (np.arange(10) + 0.5) * 20.0)
        o = pkg.rtamc2d(ctx, modes, geom, thickness=1.0, T0=300.0, dt=20.0, maxtime=200.0, nparticles=20000,
                        seed=5 + irun, sigma_bounce=2.0)
        # the CSVs hold six significant digits
        assert np.allclose(T[:, 1:], o["T"], rtol=1e-5, atol=1e-9)
        assert np.allclose(jx[:, 1:], o["jx"], rtol=1e-5, atol=1e-6 * np.abs(o["jx"]).max())


def beam_xml(nrows, path):
    """The layout of the reference's examples/example_bar/beam.xml with nrows
    rows: boxes of 50 x 2 nm stacked in y from y = 0.5, ghost boxes on both
    sides translating by -+50 nm in x, a 302 K reservoir below and a 300 K
    reservoir above."""
    H = 0.5 + 2.0 * nrows
    boxes = []
    for k in range(nrows):
        y0 = 0.5 + 2.0 * k
        boxes.append(BOX.format(i=k, x0=9.0, x1=59.0, y0=y0, y1=y0 + 2.0, teq=300.0, tin=300.0, extra=""))
    g = nrows
    boxes.append(BOX.format(i=g, x0=59.0, x1=60.5, y0=0.5, y1=H, teq=300.0, tin=300.0,
                            extra=f'\n    <Translate_to id="{g + 1}"> -50. 0.0 0.0 </Translate_to>'))
    boxes.append(BOX.format(i=g + 1, x0=7.0, x1=9.0, y0=0.5, y1=H, teq=300.0, tin=300.0,
                            extra=f'\n    <Translate_to id="{g}"> 50.0 0.0 0.0 </Translate_to>'))
    boxes.append(BOX.format(i=g + 2, x0=7.0, x1=60.5, y0=-1.5, y1=0.5, teq=302.0, tin=302.0,
                            extra="\n    <Reservoir/>"))
    boxes.append(BOX.format(i=g + 3, x0=7.0, x1=60.5, y0=H, y1=H + 2.0, teq=300.0, tin=300.0,
                            extra="\n    <Reservoir/>"))
    path.write_text(GEOM.format(n=nrows + 4, boxes="\n".join(boxes)))


def test_beam_layout_runs_in_both_solvers(pkg, tmp_path):
    beam_xml(40, tmp_path / "beam.xml")
    (tmp_path / "rta.xml").write_text(
        '<RTA_MC2d>\n<geometry file="beam.xml"/>\n<time dt="2" maxtime="200"/>\n'
        '<material name="m" modetable="synthetic" nmodes="400" seed="4" thickness="0.533" T0="300"'
        ' sigma_bounce="2.0"/>\n<particles N="100000" seed="2"/>\n</RTA_MC2d>\n')
    run([pkg.app_path("RTAMC2D"), "rta.xml", "8"], tmp_path)
    T = np.loadtxt(tmp_path / "temperature_300K_run_0.csv", delimiter=",")
    jy = np.loadtxt(tmp_path / "jy_300K_run_0.csv", delimiter=",")
    assert T.shape == (100, 1 + 44)
    # heat flows up from the 302 K reservoir: positive flux, T falling with y
    assert (jy[-1, 1:41] > 0).all()
    assert T[-1, 1:6].mean() > T[-1, 36:41].mean()
    run([pkg.app_path("PropagatorBuilder"), "synthetic:400:4", "300", "0.5", "1", "20"], tmp_path)
    (tmp_path / "berta.xml").write_text(
        '<beRTAMC2D>\n<geometry file="beam.xml"/>\n<time dt="0.5" maxtime="20"/>\n'
        '<material name="m" modetable="synthetic" nmodes="400" seed="4" thickness="0.533">\n'
        '  <propagator T="300" file="synthetic_400_4_300K_0.5ps.P.eigen.bin"/>\n</material>\n'
        '<Eeff Ed="5.39836e-25"/>\n</beRTAMC2D>\n')
    r = run([pkg.app_path("beRTAMC2D"), "berta.xml", "4"], tmp_path)
    lines = [ln.split("\t") for ln in r.stdout.splitlines() if ln and not ln.startswith("#")]
    rows = np.array([[float(x) for x in ln if x != ""] for ln in lines])
    assert rows.shape[1] == 3 + 3 * 44
    # particles enter from the hot reservoir and energy appears in the first rows
    assert rows[-1, 2] > 0 and rows[-1, 3] > 0


def test_rtamc2d_steady_csv(pkg, tmp_path):
    """<gradient> selects the steady branch: steady_<T0>K_run_<i>.csv holds
    the rows T, jx, jy over the boxes."""
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(80.0, 30.0, 8, 300.0, reservoirs=(300.0 + 0.08, 300.0 - 0.08))
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, teq)
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d>\n<geometry file="g.xml"/>\n<time dt="20" maxtime="200"/>\n<gradient x="-2.0e-3" y="0.0"/>\n'
        '<convergence energy="1e-3" flux="1e-3"/>\n'
        '<material name="m" modetable="synthetic" nmodes="300" seed="5" thickness="1" T0="300" sigma_bounce="2.0"/>\n'
        '<particles N="50000" seed="3"/>\n</RTA_MC2d>\n')
    r = run([pkg.app_path("RTAMC2D"), "in.xml", "8"], tmp_path)
    assert "generations" in r.stdout
    out = np.loadtxt(tmp_path / "steady_300K_run_0.csv", delimiter=",")
    assert out.shape == (3, 10)
    assert (out[1, :8] > 0).all()  # heat flows down the imposed gradient
    assert np.abs(out[0, :8] - 300.0).max() < 1.0


def test_reference_example_bar(pkg, tmp_path):
    """The reference's examples/example_bar inputs (tests/golden/
    example_bar_beam.xml is its beam.xml: 200 boxes of 50 x 2 nm, periodic
    ghost boxes on both sides, a 302 K reservoir below and a 300 K one
    above) through RTAMC2D with <spectral>, beRTAMC2D and dist_reader, with
    a synthetic mode table in place of the HDF5 database and shorter runs."""
    import shutil
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_bar_beam.xml")
    shutil.copy(golden, tmp_path / "beam.xml")
    mat = '<material name="black_P" modetable="synthetic" nmodes="400" seed="4" thickness="0.533" T0="300.0"'
    (tmp_path / "rta.xml").write_text(
        '<RTA_MC2d>\n<geometry file="beam.xml"/>\n<time dt="0.25" maxtime="100"/>\n' + mat +
        ' sigma_bounce="2.0"/>\n<particles N="200000"/>\n<spectral>\n  <resolution ticks="50"/>\n'
        '  <location bin="5"/>\n  <location bin="99"/>\n</spectral>\n</RTA_MC2d>\n')
    run([pkg.app_path("RTAMC2D"), "rta.xml", "8"], tmp_path)
    T = np.loadtxt(tmp_path / "temperature_300K_run_0.csv", delimiter=",")
    jy = np.loadtxt(tmp_path / "jy_300K_run_0.csv", delimiter=",")
    assert T.shape == (400, 1 + 204)
    # heat enters from the 302 K reservoir (box 202) below: the lowest rows
    # warm first and the flux points up
    assert T[-1, 1:11].mean() > T[-1, 191:201].mean()
    assert jy[-40:, 1:11].mean() > 0
    sp = np.loadtxt(tmp_path / "deltaT_omega_5_300K_run_0.csv", delimiter=",")
    assert sp.shape == (401, 51)
    run([pkg.app_path("PropagatorBuilder"), "synthetic:400:4", "300", "0.25", "1", "20"], tmp_path)
    (tmp_path / "berta.xml").write_text(
        '<beRTAMC2D>\n<geometry file="beam.xml"/>\n<time dt="0.25" maxtime="10"/>\n' + mat + '>\n'
        '  <propagator T="300" file="synthetic_400_4_300K_0.25ps.P.eigen.bin"/>\n</material>\n'
        '<Eeff Ed="5.39836e-25"/>\n<spectral>\n  <resolution ticks="50"/>\n  <time t="2."/>\n  <time t="5."/>\n'
        '  <location bin="5"/>\n  <location bin="6"/>\n</spectral>\n</beRTAMC2D>\n')
    r = run([pkg.app_path("beRTAMC2D"), "berta.xml", "4"], tmp_path)
    lines = [ln.split("\t") for ln in r.stdout.splitlines() if ln and not ln.startswith("#")]
    rows = np.array([[float(x) for x in ln if x != ""] for ln in lines])
    assert rows.shape[1] == 3 + 3 * 204 and rows.shape[0] >= 41
    assert rows[-1, 2] > 0 and rows[-1, 3] > 0  # particles from the hot reservoir in box 0
    run([pkg.app_path("dist_reader"), "berta.xml", "properties.msgpack.bin"], tmp_path)
    d = np.loadtxt(tmp_path / "deltaT_omega_5.csv", delimiter=",")
    assert d.shape == (3, 51) and d[1, 0] == 2.0 and d[2, 0] == 5.0


def test_rtamc2d_reads_per_mode_sigma_from_the_mode_file(pkg, ctx, tmp_path):
    """A .modes file with a fifth column gives the per-mode bounce broadening
    (base_sigma): RTAMC2D then writes what the library gives with
    ModeTable.set_sigma on the same seeds, and something else than with the
    constant sigma_bounce."""
    W = pkg.workloads
    rect, kind, teq, _ = W.ribbon(60.0, 20.0, 6, 300.0, reservoirs=(300.5, 299.5))
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, teq)
    mat = W.material(300, seed=4)
    sig = 0.2 + 0.3 * np.hypot(mat["vx"], mat["vy"])
    rows = "\n".join(" ".join(repr(float(v)) for v in r) for r in
                     zip(mat["omega"], mat["vx"], mat["vy"], mat["w"], sig))
    (tmp_path / "m.modes").write_text(f"300 1 {float(mat['V'])!r} {float(mat['lattz'])!r}\n{rows}\n")
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d>\n<geometry file="g.xml"/>\n<time dt="20" maxtime="200"/>\n'
        '<material name="m" modetable="m.modes" thickness="1" T0="300" sigma_bounce="2.0"/>\n'
        '<particles N="20000" seed="5"/>\n</RTA_MC2d>\n')
    run([pkg.app_path("RTAMC2D"), "in.xml", "8", "1"], tmp_path)
    T = np.loadtxt(tmp_path / "temperature_300K_run_0.csv", delimiter=",")
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, rect, kind, teq)
    kw = dict(thickness=1.0, T0=300.0, dt=20.0, maxtime=200.0, nparticles=20000, seed=5, sigma_bounce=2.0)
    const = pkg.rtamc2d(ctx, modes, geom, **kw)
    modes.set_sigma(sig)
    o = pkg.rtamc2d(ctx, modes, geom, **kw)
    assert np.allclose(T[:, 1:], o["T"], rtol=1e-5, atol=1e-9)
    assert not np.allclose(o["T"], const["T"], rtol=0, atol=1e-9)


def test_bertamc2d_reads_init_distribution(pkg, ctx, tmp_path):
    """<init_distribution box file/>: column 0 of a tab-separated file,
    '#' lines skipped (get_column, beyondRTA_MC2d.cpp:1270-1291). The
    executable starts from the same particles as the library run with the
    same distribution and seed (equal initial count and properties)."""
    W = pkg.workloads
    rect, kind, teq, tin = W.ribbon(60.0, 30.0, 6, 300.0)
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, tin)
    mat = W.material(200, seed=2)
    n = len(mat["omega"])
    fd = 1e-23 * (1.0 + np.arange(n) % 7) * np.where(np.arange(n) % 5 == 0, -1.0, 1.0)
    (tmp_path / "fd.dat").write_text("# given distribution\n" + "\n".join(f"{float(v)!r}\t0" for v in fd) + "\n")
    (tmp_path / "in.xml").write_text(
        '<beRTAMC2D>\n<geometry file="g.xml"/>\n<time dt="5" maxtime="20"/>\n'
        '<material name="m" modetable="synthetic" nmodes="200" seed="2" thickness="1" T0="300">\n'
        '  <propagator T="300" file="synthetic" density="0.05" seed="1" m="20"/>\n</material>\n'
        '<Eeff Ed="1e-22"/>\n<init_distribution box="2" file="fd.dat"/>\n<seed value="9"/>\n</beRTAMC2D>\n')
    r = run([pkg.app_path("beRTAMC2D"), "in.xml", "2"], tmp_path)
    assert "init_distribution" in r.stdout
    first = [l for l in r.stdout.splitlines() if l and not l.startswith("#")][0].split("\t")
    assert int(first[2]) > 100   # particles at step 0


def test_rtamc2d_runs_polygon_geometries(pkg, ctx, tmp_path):
    """A geometry file with quadrilateral boxes (slanted shared sides and a
    zigzag top wall) between two rectangular reservoirs: RTAMC2D writes what
    the library's polygon geometry gives on the same seed."""
    from test_gpu_polygons import zigzag_ribbon
    polys = zigzag_ribbon(6, 60.0, 30.0, 3.0, 4.0, reservoirs=True)
    kind = np.array([0] * 6 + [1, 1], np.int32)
    teq = np.array([300.0] * 6 + [300.5, 299.5])
    boxes = []
    for i, p in enumerate(polys):
        pts = " ".join(f"{float(x)!r} {float(y)!r}" for x, y in p)
        res = "\n    <Reservoir/>" if kind[i] == 1 else ""
        boxes.append(f'  <Box>\n    <BoxID id="{i}"/>\n    <MaterialID name="m"/>\n'
                     f'    <Vertices dim="2" npoints="{len(p)}"> {pts} </Vertices>\n'
                     f'    <initCnd Teq="{float(teq[i])!r}"/>{res}\n  </Box>')
    (tmp_path / "g.xml").write_text(f'<Geometry>\n  <number_of_boxes Ngeom="{len(polys)}"/>\n' +
                                    "\n".join(boxes) + "\n</Geometry>\n")
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d>\n<geometry file="g.xml"/>\n<time dt="20" maxtime="200"/>\n'
        '<material name="m" modetable="synthetic" nmodes="300" seed="4" thickness="1" T0="300" sigma_bounce="1.0"/>\n'
        '<particles N="20000" seed="5"/>\n</RTA_MC2d>\n')
    run([pkg.app_path("RTAMC2D"), "in.xml", "8", "1"], tmp_path)
    T = np.loadtxt(tmp_path / "temperature_300K_run_0.csv", delimiter=",")
    mat = pkg.workloads.material(300, seed=4)
    modes = pkg.ModeTable(ctx, mat["omega"], mat["vx"], mat["vy"], mat["w"], mat["V"], mat["lattz"])
    geom = pkg.Geometry(ctx, None, kind, teq, polys=polys)
    o = pkg.rtamc2d(ctx, modes, geom, thickness=1.0, T0=300.0, dt=20.0, maxtime=200.0, nparticles=20000, seed=5,
                    sigma_bounce=1.0)
    assert np.allclose(T[:, 1:], o["T"], rtol=1e-5, atol=1e-9)
    assert np.abs(o["T"][:, :6] - 300.0).max() > 1e-3
