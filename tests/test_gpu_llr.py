"""The device LLR (energy-window constrained sweeps, drive-in, Robbins-Monro):
bit-exact against the oracle's restatement of the same algorithm, the same
result alone, in a mixed engine and through the serving loop, and a_n in
statistical agreement with the reference llr binary (tests/golden/llr_stats_4x4.txt)."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

KEYS = ["a", "S1", "S2", "last_da", "maxdrift", "phase", "iterations", "seed_sweeps", "reached", "sweeps", "hit",
        "acc"]

# group, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start, K, NRM
CASES = [
    ("mys1080-v4", 4, 4, 4, 1, -0.1, -1500.0, 75.0, 1.5, 3, "hot", 5, 4),
    ("mys1080-v4", 4, 4, 4, 1, -0.1, -3000.0, 60.0, 2.5, 4, "hot", 4, 3),
    ("S216", 4, 4, 6, 2, 1.0, -1200.0, 80.0, 0.5, 5, "hot", 4, 3),
    ("mys1080-v4", 4, 6, 4, 1, 0.0, -2000.0, 100.0, 1.8, 6, "cold", 4, 3),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(dsa, case):
    name, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start, K, NRM = case
    return dsa.dsa.llr_params(D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start=start, K=K, NRM=NRM)


def oracle_chain(oracle, ogroups, case):
    name, D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start, K, NRM = case
    return oracle.OLlr(ogroups(name), D, Nt, Nx, axis, bg, Ewin, hw, a0, seed, start=start, K=K, NRM=NRM)


def same(gpu, ora):
    for k in KEYS:
        assert gpu[k] == ora[k], (k, gpu[k], ora[k])


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[2]}x{c[3]}-axis{c[4]}-{c[10]}")
def test_llr_matches_oracle(dsa, dgroups, oracle, ogroups, case):
    say("case llr vs oracle", case)
    o = oracle_chain(oracle, ogroups, case)
    with dsa.dsa.LlrEngine(4, 4) as eng:
        s = eng.add(dgroups(case[0]), params(dsa, case))
        for sweep in range(30):
            eng.step(1)
            o.step(1)
            same(eng.status(s), o.status())
            assert np.array_equal(eng.links(s), o.a_links), f"links differ after sweep {sweep + 1}"


def test_llr_mix_equals_alone(dsa, dgroups, oracle, ogroups):
    say("case llr mix")
    orc = [oracle_chain(oracle, ogroups, c) for c in CASES]
    with dsa.dsa.LlrEngine(4, 8) as eng:
        slots = [eng.add(dgroups(c[0]), params(dsa, c)) for c in CASES]
        for _ in range(3):
            eng.step(8)
            for o in orc:
                o.step(8)
            for s, o in zip(slots, orc):
                same(eng.status(s), o.status())
                assert np.array_equal(eng.links(s), o.a_links)


def test_llr_serve_results_equal_alone(dsa, dgroups, oracle, ogroups):
    """Requests arriving while others run are admitted as slots free up (3
    slots for 6 requests); each retires with exactly its lone result."""
    say("case llr serve")
    reqs = CASES + [CASES[0][:9] + (11,) + CASES[0][10:], CASES[2][:9] + (12,) + CASES[2][10:]]
    want = []
    for c in reqs:
        o = oracle_chain(oracle, ogroups, c)
        for _ in range(2000):
            o.step(1)
            if o.status()["phase"] == 2:
                break
        want.append(o.status())
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.03, 0.05]
    with dsa.dsa.LlrEngine(4, 3) as eng:
        admit, done, st = eng.serve([dgroups(c[0]) for c in reqs], [params(dsa, c) for c in reqs], arrivals, chunk=3)
    for r in range(len(reqs)):
        same(st[r], want[r])
        assert admit[r] >= arrivals[r] and done[r] >= admit[r]


def test_llr_statistics_match_reference_llr(dsa, dgroups):
    """Window S1 in [-1575, -1425] on 4^4, bg = -0.1, a0 = 1.5, K = 20, NRM = 30:
    16 chains (8 hot, 8 cold starts, like llr's R = 8 down/up passes).  The
    mean a_n agrees with the reference llr's within 2 standard errors."""
    say("case llr statistics")
    ref = [l for l in open(os.path.join(GOLDEN, "llr_stats_4x4.txt")) if l.startswith("ANE:")][0].split()
    m_ref, sd_ref, n_ref = float(ref[2]), float(ref[3]), int(ref[4])
    with dsa.dsa.LlrEngine(4, 16) as eng:
        slots = []
        for k in range(16):
            p = dsa.dsa.llr_params(4, 4, 4, 1, -0.1, -1500.0, 75.0, 1.5, 700 + k, start="hot" if k < 8 else "cold",
                                   K=20, NRM=30)
            slots.append(eng.add(dgroups("mys1080-v4"), p))
        for _ in range(200):
            eng.step(5)
            if all(eng.status(s)["phase"] == 2 for s in slots):
                break
        st = [eng.status(s) for s in slots]
    assert all(s["reached"] == 1 and s["iterations"] == 30 for s in st)
    a = np.array([s["a"] for s in st])
    se = np.hypot(sd_ref / np.sqrt(n_ref), a.std(ddof=1) / np.sqrt(len(a)))
    z = abs(a.mean() - m_ref) / se
    say(f"  a_n: reference {m_ref:.5f} +- {sd_ref:.5f} (n={n_ref}); device {a.mean():.5f} +- {a.std(ddof=1):.5f}; "
        f"z = {z:.2f}")
    assert z < 2.0
