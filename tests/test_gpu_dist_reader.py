"""dist_reader on the GPU (spectral post-processing of beRTAMC2D frames,
src/dist_reader.cpp:271-508) against a numpy computation from the decoded
frames: Gamma densities from scipy, the reference's tick mesh and
normalisation."""
import numpy as np
import pytest
import scipy.stats

from test_gpu_frontends import eigen_bin, run, write_ribbon

pytestmark = pytest.mark.gpu

KB, HBAR = 1.3806488e-23, 1.054571726e-34


def be_kernel(om, T):
    x = 0.5e12 * HBAR * om / (KB * T)
    return (x / np.sinh(x)) ** 2


def test_dist_reader_spectra(pkg, tmp_path):
    msgpack = pytest.importorskip("msgpack")
    W = pkg.workloads
    run([pkg.app_path("PropagatorBuilder"), "synthetic:96:2", "300", "5", "1", "20"], tmp_path)
    rect, kind, teq, tin = W.ribbon(60.0, 30.0, 6, 300.0, Tinit_first=301.0)
    write_ribbon(tmp_path / "g.xml", rect, kind, teq, tin)
    (tmp_path / "in.xml").write_text(
        '<beRTAMC2D>\n<geometry file="g.xml"/>\n<time dt="5" maxtime="45"/>\n'
        '<material name="m" modetable="synthetic" nmodes="96" seed="2" thickness="1" T0="300">\n'
        '  <propagator T="300" file="synthetic_96_2_300K_5ps.P.eigen.bin"/>\n</material>\n'
        '<Eeff particles="5000" Tmax="301" Tmin="300"/>\n<seed value="7"/>\n'
        '<spectral>\n  <resolution ticks="30"/>\n  <time t="25."/>\n  <time t="10."/>\n'
        '  <location bin="0"/>\n  <location bin="2"/>\n</spectral>\n</beRTAMC2D>\n')
    run([pkg.app_path("beRTAMC2D"), "in.xml", "4"], tmp_path)
    run([pkg.app_path("dist_reader"), "in.xml", "properties.msgpack.bin"], tmp_path)
    # frames
    raw = (tmp_path / "properties.msgpack.bin").read_bytes()
    frames, pos = {}, 0
    while True:
        k = raw.index(b"#", pos)
        size = int(raw[pos:k])
        if size == 0:
            break
        f = msgpack.unpackb(raw[k + 1:k + 1 + size])
        frames[round(f[0], 6)] = f
        pos = k + 1 + size
    mat = W.material(96, seed=2)
    om, vx, vy = mat["omega"], mat["vx"], mat["vy"]
    K, sigma = 30, 0.25
    mesh = np.linspace(0.0, om.max(), K)
    mesh += mesh[1] / 2.0
    live = om > 1e-8
    w = np.zeros((96, K))
    w[live] = scipy.stats.gamma.pdf(mesh[None, :], a=(om[live] ** 2 / sigma)[:, None],
                                    scale=(sigma / om[live])[:, None])
    vuc = mat["V"] * 1.0 / mat["lattz"]
    cvtot = KB / vuc * be_kernel(om[live], 300.0).sum()
    for box in (0, 2):
        vol = (rect[box, 1] - rect[box, 0]) * (rect[box, 3] - rect[box, 2]) * 1.0
        files = {name: np.loadtxt(tmp_path / f"{name}_{box}.csv", delimiter=",")
                 for name in ("deltaT_omega", "jx_omega", "jy_omega")}
        fd = np.loadtxt(tmp_path / f"fd_{box}.csv", delimiter=",", comments="#")
        assert (tmp_path / f"fd_{box}.csv").read_text().startswith("# 0\n# 0\n")
        for a in files.values():
            assert a.shape == (3, K + 1) and a[0, 0] == -1.0
            assert np.allclose(a[0, 1:], mesh, rtol=1e-5)
        for row, t in enumerate((10.0, 25.0)):
            f = frames[t]
            Eeff = f[1]
            h = np.zeros(96)
            pairs = f[4 + 2 * box]
            if f[3 + 2 * box]:
                h[np.array(pairs[0::2])] = pairs[1::2]
            d = h @ w * Eeff / (vol * cvtot)
            jx = (h * vx) @ w * Eeff * 1e30 / vol
            jy = (h * vy) @ w * Eeff * 1e30 / vol
            fq = (h[live] / om[live]).sum() * Eeff * vuc / (1e12 * vol * HBAR)
            for name, ref in (("deltaT_omega", d), ("jx_omega", jx), ("jy_omega", jy)):
                got = files[name][row + 1]
                assert got[0] == t
                assert np.allclose(got[1:], ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max()), (box, t, name)
            assert fd[row, 0] == t and fd[row, 1] == pytest.approx(fq, rel=2e-5)
        assert np.abs(files["deltaT_omega"][1:, 1:]).max() > 0
