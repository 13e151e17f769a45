"""The oracle's PCG32 against the reference's own known-answer file
(external/pcg-cpp-0.98/test-high/expected/check-pcg32.out, produced by
test-high/pcg-test.cpp:80-180 with pcg32 rng(42u, 54u)): all five rounds of
raw outputs, coin tosses, dice rolls and the card shuffle."""
import os

import oracle_lib as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "check-pcg32.out")
NUMBER = "A23456789TJQK"
SUIT = "hcds"


def render_rounds(rounds=5):
    rng = orc.Pcg32(42, 54)
    lines = []
    for rnd in range(1, rounds + 1):
        lines.append(f"Round {rnd}:")
        vals = rng.next(6)
        lines.append("  32bit: " + " ".join(f"0x{v:08x}" for v in vals))
        # "Again" re-generates the same six outputs after backstep(6)
        lines.append("  Again: " + " ".join(f"0x{v:08x}" for v in vals))
        lines.append("  Coins: " + "".join("H" if c else "T" for c in rng.bounded(2, 65)))
        st0 = rng.st[0]
        rolls = rng.bounded(6, 33)
        # each bounded draw with bound 6 consumes one output unless rejected
        lines.append("  Rolls: " + " ".join(str(r + 1) for r in rolls))
        used = _distance(st0, rng.st[0], rng.st[1])
        lines.append(f"   -->   rolling dice used {used} random numbers")
        cards = list(range(52))
        count = 52
        while count > 1:
            chosen = rng.bounded(count)[0]
            count -= 1
            cards[chosen], cards[count] = cards[count], cards[chosen]
        txt = [f"{NUMBER[c // 4]}{SUIT[c % 4]}" for c in cards]
        lines.append("  Cards: " + " ".join(txt[:22]))
        lines.append("\t " + " ".join(txt[22:44]))
        lines.append("\t " + " ".join(txt[44:]))
        lines.append("")
    return lines


def _distance(s0, s1, inc):
    """number of LCG steps from state s0 to s1 (brute force, small)"""
    mult = 6364136223846793005
    m = (1 << 64) - 1
    s = s0
    for k in range(1000):
        if s == s1:
            return k
        s = (s * mult + inc) & m
    raise AssertionError("states too far apart")


def test_pcg32_known_answers():
    golden = open(GOLDEN).read().splitlines()
    start = golden.index("Round 1:")
    expect = [ln.rstrip() for ln in golden[start:]]
    got = [ln.rstrip() for ln in render_rounds(5)]
    assert len(got) == 50 and len(expect) >= 49
    assert got[: len(expect)] == expect[: len(got)]
    assert len([ln for ln in expect if ln.startswith("Round")]) == 5


def test_uniform_is_53_bit_and_in_range():
    rng = orc.Pcg32(7, 11)
    u = rng.uniform(20000)
    assert (u >= 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 0.01
    # exactly representable on the 2^-53 grid
    assert ((u * 2.0 ** 53) % 1 == 0).all()


def test_bounded_is_uniform():
    rng = orc.Pcg32(3, 5)
    import numpy as np
    x = np.array(rng.bounded(7, 70000))
    counts = np.bincount(x, minlength=7)
    chi2 = ((counts - 10000) ** 2 / 10000).sum()
    assert chi2 < 30  # 6 dof, p ~ 5e-5
