"""Front-end executables (RTAMC2D, beRTAMC2D, PropagatorBuilder) without a
GPU: input validation in the reference's XML schemas, the msgpack frame
encoding (beyondRTA_MC2d.cpp:1612-1769) decoded with the msgpack library,
and the .eigen.bin layout (collision_operator.cpp:931-1013)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def app(pkg, name):
    return pkg.app_path(name)


def geometry_xml(boxes):
    out = ["<Geometry>", f'  <number_of_boxes Ngeom="{len(boxes)}"/>']
    for i, b in enumerate(boxes):
        (x0, x1, y0, y1), extra = b[0], b[1] if len(b) > 1 else ""
        out += ["  <Box>", f'    <BoxID id="{i}"/>', '    <MaterialID name="m"/>',
                '    <Vertices dim="2" npoints="4">', f"      {x0} {y0} {x0} {y1} {x1} {y1} {x1} {y0}",
                "    </Vertices>", '    <initCnd Teq="300."/>', extra, "  </Box>"]
    out.append("</Geometry>")
    return "\n".join(out)


def run(cmd, cwd):
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_usage_errors(pkg, tmp_path):
    for name in ("RTAMC2D", "beRTAMC2D", "PropagatorBuilder"):
        r = run([app(pkg, name)], tmp_path)
        assert r.returncode == 1 and "Usage" in r.stderr
    r = run([app(pkg, "dist_reader")], tmp_path)  # dist_reader.cpp:547-551
    assert r.returncode == 1 and "USAGE: dist_reader" in r.stdout


def test_rtamc2d_rejects_hdf5_material(pkg, tmp_path):
    (tmp_path / "g.xml").write_text(geometry_xml([((0, 10, 0, 5),)]))
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d><geometry file="g.xml"/><time dt="0.25" maxtime="10"/>'
        '<material name="m" database="x.h5" thickness="1" T0="300"/><particles N="10"/></RTA_MC2d>')
    r = run([app(pkg, "RTAMC2D"), "in.xml", "4"], tmp_path)
    assert r.returncode == 1 and "HDF5" in r.stderr


@pytest.mark.parametrize("extra,msg", [
    ('<Translate_to id="0"> 1.0 </Translate_to>', "bad <Translate_to>"),
])
def test_geometry_features_rejected(pkg, tmp_path, extra, msg):
    (tmp_path / "g.xml").write_text(geometry_xml([((0, 10, 0, 5), extra)]))
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d><geometry file="g.xml"/><time dt="0.25" maxtime="10"/>'
        '<material name="m" modetable="synthetic" nmodes="50" thickness="1" T0="300"/>'
        '<particles N="10"/></RTA_MC2d>')
    r = run([app(pkg, "RTAMC2D"), "in.xml", "4"], tmp_path)
    assert r.returncode == 1 and msg in r.stderr


def test_theta_is_read_and_not_applied(pkg, tmp_path):
    """<theta angle> sets geometry_2d::rotmat, which neither solver reads
    (geometry_2d.cpp:496-509): the reader accepts it and says so"""
    (tmp_path / "g.xml").write_text(geometry_xml([((0, 10, 0, 5), '<theta angle="0.3"/>')]))
    (tmp_path / "in.xml").write_text(
        '<RTA_MC2d><geometry file="g.xml"/><time dt="0.25" maxtime="10"/>'
        '<material name="m" modetable="synthetic" nmodes="50" thickness="1" T0="300"/>'
        '<particles N="10"/></RTA_MC2d>')
    r = run([app(pkg, "RTAMC2D"), "in.xml", "4"], tmp_path)
    assert "theta 0.3 is read and, as in the reference, not applied" in r.stderr
    assert "rotated" not in r.stderr


def test_xml_errors_name_the_line(pkg, tmp_path):
    (tmp_path / "in.xml").write_text('<RTA_MC2d>\n<geometry file="g.xml">\n</RTA_MC2d>\n')
    r = run([app(pkg, "RTAMC2D"), "in.xml", "4"], tmp_path)
    assert r.returncode == 1 and "line 3" in r.stderr


def test_propagator_builder_validates_arguments(pkg, tmp_path):
    r = run([app(pkg, "PropagatorBuilder"), "synthetic:8:1", "300", "5", "4", "10", "fp16"], tmp_path)
    assert r.returncode == 1 and "precision" in r.stderr
    r = run([app(pkg, "PropagatorBuilder"), "missing.B.eigen.bin", "300", "5"], tmp_path)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_frame_and_eigen_bin_encoding(pkg, tmp_path):
    msgpack = pytest.importorskip("msgpack")
    exe = tmp_path / "io_check"
    lib = os.path.join(ROOT, "bte-barna_amd", "lib")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "bte-barna_amd", "apps"), os.path.join(ROOT, "tests", "frontend_io_check.cpp"),
                    "-o", str(exe), "-L", lib, "-lbte_amd", f"-Wl,-rpath,{lib}"], check=True, timeout=120)
    r = run([str(exe), "frames.bin", "m.eigen.bin"], tmp_path)
    assert r.returncode == 0
    raw = (tmp_path / "frames.bin").read_bytes()
    frames, pos = [], 0
    while True:
        k = raw.index(b"#", pos)
        size = int(raw[pos:k])
        if size == 0:
            assert k + 1 == len(raw)
            break
        frames.append(msgpack.unpackb(raw[k + 1:k + 1 + size], strict_map_key=False))
        pos = k + 1 + size
    assert len(frames) == 2
    f = frames[0]
    assert len(f) == 3 + 2 * 3
    assert f[0] == 12.5 and f[1] == 3.25e-24 and f[2] == [300.0, 301.5, 299.25]
    assert f[3] == 8 and f[4] == [1, 1, 2, -1, 3, 127, 4, 128]
    assert f[5] == 0 and f[6] == [42.0]
    assert f[7] == 10 and f[8] == [0, -32, 1, -33, 2, 70000, 3, -129, 4, -40000]
    assert frames[1][0] == 17.5
    b = (tmp_path / "m.eigen.bin").read_bytes()
    rows = np.frombuffer(b[:8], np.int64)[0]
    a = np.frombuffer(b[8:], np.float64).reshape(3, 3, order="F")
    assert rows == 3 and len(b) == 8 + 9 * 8
    assert a[1, 0] == -0.5 and a[0, 1] == 0.5  # column-major: element (i, j) at i + 3 j


def test_krylov_size_matches_oracle(pkg):
    import oracle_lib as orc
    mat, B = pkg.workloads.config1_inputs(nmodes=64, seed=3)
    for dt in (0.05, 0.5, 2.0):
        assert pkg.krylov_size(B, dt) == orc.krylov_size(B, dt, 1e-8)
