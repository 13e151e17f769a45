"""The C ABI library (no GPU needed): every declared entry point is
exported and bound, and the host-side logic behind it (synthetic inputs,
alias tables, argument checks) is exact."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bte_abi.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(bte_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg._abi.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert sorted(pkg._abi.SIGNATURES) == names


def test_missing_library_fails_loudly(pkg, monkeypatch):
    A = pkg._abi
    monkeypatch.setattr(A, "_lib", None)
    monkeypatch.setattr(A, "LIBPATH", os.path.join(ROOT, "no_such_lib.so"))
    with pytest.raises(RuntimeError, match="missing"):
        A.lib()


def test_version(pkg):
    assert pkg._abi.lib().bte_version() >= 100


def test_synthetic_modes_match_the_oracle(pkg):
    for mirror in (False, True):
        a = pkg.synthetic_modes(1000, 17, mirror=mirror)
        b = orc.synthetic_modes(1000, 17, mirror=mirror)
        for k in ("omega", "w"):
            assert np.array_equal(a[k], b[k]), k
        for k in ("vx", "vy"):  # host libm sin/cos may differ by one ulp
            np.testing.assert_allclose(a[k], b[k], rtol=4e-16, atol=4e-16)
    m = pkg.synthetic_modes(4000, 1)
    f = m["omega"] / (2 * np.pi)
    assert f.min() >= 0.1 and f.max() <= 15.0
    sp = np.hypot(m["vx"], m["vy"])
    assert sp.min() >= 0.5 and sp.max() <= 8.0
    tau = 1 / m["w"]
    assert tau.min() >= 1.0 and tau.max() <= 1000.0


def test_mirror_table_pairs(pkg):
    m = pkg.synthetic_modes(10, 3, mirror=True)
    np.testing.assert_array_equal(m["vy"][5:], -m["vy"][:5])
    np.testing.assert_array_equal(m["vx"][5:], m["vx"][:5])


def test_synthetic_B_matches_the_oracle(pkg):
    m = orc.synthetic_modes(300, 2)
    Cm = orc.bose_einstein_kernel(m["omega"], 300.0)
    a = pkg.synthetic_B(Cm, m["w"], density=0.1, seed=4)
    b = orc.synthetic_B(Cm, m["w"], 0.1, 4)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [1, 2, 7, 100, 2000])
def test_alias_tables_are_exact(pkg, n):
    rng = np.random.default_rng(n)
    p = rng.random(n) ** 3
    p[rng.random(n) < 0.2] = 0.0
    if p.sum() == 0:
        p[0] = 1.0
    th, al = pkg.alias_build(p)
    # probability of k: own column accepted + alias mass pointing at k
    acc = th.astype(np.float64) / 2.0 ** 32
    acc[th == 0xFFFFFFFF] = 1.0
    prob = acc.copy()
    np.add.at(prob, al, 1.0 - acc)
    np.testing.assert_allclose(prob / n, p / p.sum(), atol=4 / 2.0 ** 32)
    assert ((p == 0) <= (prob < 1e-9)).all()


def test_argument_errors_are_reported(pkg):
    L = pkg._abi.lib()
    rc = L.bte_synthetic_modes(0, 1, 0.1, 15.0, 0.5, 8.0, 1.0, 1000.0, 0, None, None, None, None)
    assert rc == -1 and "bad arguments" in pkg._abi.last_error()
    with pytest.raises(pkg.BteError, match="EINVAL"):
        pkg.alias_build(np.array([1.0, -1.0]))
    with pytest.raises(pkg.BteError, match="EINVAL"):
        pkg.alias_build(np.zeros(4))
