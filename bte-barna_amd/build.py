"""Builds the MI355X engine: libbte_amd.so (HIP kernels + C ABI) and the
RTAMC2D / beRTAMC2D / PropagatorBuilder front-ends, in-tree under lib/."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
APPS = os.path.join(HERE, "apps")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BTE_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument",
]
SOURCES = ["host_util.cpp", "rta.hip", "bert.hip", "krylov.hip"]
APP_SOURCES = ["RTAMC2D.cpp", "beRTAMC2D.cpp", "PropagatorBuilder.cpp", "dist_reader.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
    _run([HIPCC, *CFLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", out])
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    so = os.path.join(LIB, "libbte_amd.so")
    _run([HIPCC, *CFLAGS, "-shared", "-o", so + ".tmp", *objs, "-lrocblas", "-lrocsolver"])
    os.replace(so + ".tmp", so)
    if os.path.isdir(APPS):
        for app in APP_SOURCES:
            src = os.path.join(APPS, app)
            if not os.path.exists(src):
                continue
            exe = os.path.join(LIB, os.path.splitext(app)[0])
            _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe + ".tmp",
                  "-L", LIB, "-lbte_amd", f"-Wl,-rpath,$ORIGIN", "-lpthread"])
            os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", so)
    return so


if __name__ == "__main__":
    build(verbose=True)
