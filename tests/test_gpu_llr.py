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

KEYS = ["a", "S1", "S2", "last_da", "maxdrift", "drift1", "drift2", "phase", "iterations", "seed_sweeps", "reached",
        "sweeps", "hit", "acc"]

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
@pytest.mark.parametrize("kernel", ["table", "class", "class-all-exact", "field", "field-all-exact"])
def test_llr_matches_oracle(dsa, dgroups, oracle, ogroups, case, kernel):
    """Every pass kernel (element values; class values with certified
    decisions, with the staple-times-link products from table gathers or
    from the GF(4) form, and with every hit sent to its exact path) against
    the oracle."""
    say("case llr vs oracle", case, kernel)
    o = oracle_chain(oracle, ogroups, case)
    with dsa.dsa.LlrEngine(4, 4) as eng:
        if kernel.endswith("-all-exact"):
            eng.set_kernel(kernel[: -len("-all-exact")], 1e300)
        else:
            eng.set_kernel(kernel)
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
    mean a_n agrees with the reference llr's (Welch's t-test, two-sided
    2-sigma level)."""
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
    from stats_util import p_limit, welch_p

    p = welch_p(a.mean(), a.std(ddof=1), a.size, m_ref, sd_ref, n_ref)
    say(f"  a_n: reference {m_ref:.5f} +- {sd_ref:.5f} (n={n_ref}); device {a.mean():.5f} +- {a.std(ddof=1):.5f}; "
        f"p = {p:.3f}")
    assert p > p_limit(1)


# axis 3: box (Ewin, hw) x (Ewin2, hw2) on (S1, S2), llr2d.cpp
CASES2D = [
    # group, D, Nt, Nx, E1, hw1, E2, hw2, a0, a20, seed, start, K, NRM
    ("mys1080-v4", 4, 3, 3, -500.0, 40.0, -400.0, 40.0, 1.0, 1.0, 21, "hot", 5, 4),
    ("mys1080-v4", 4, 4, 4, -1200.0, 90.0, -1300.0, 120.0, 1.6, 0.3, 22, "cold", 4, 3),
    ("S216", 4, 4, 4, -900.0, 80.0, -700.0, 100.0, 1.0, 0.5, 23, "hot", 4, 3),
]


def params2d(dsa, c):
    name, D, Nt, Nx, E1, hw1, E2, hw2, a0, a20, seed, start, K, NRM = c
    return dsa.dsa.llr_params(D, Nt, Nx, 3, 0.0, E1, hw1, a0, seed, start=start, K=K, NRM=NRM, Ewin2=E2, hw2=hw2,
                              a20=a20)


def oracle2d(oracle, ogroups, c, links=None):
    name, D, Nt, Nx, E1, hw1, E2, hw2, a0, a20, seed, start, K, NRM = c
    return oracle.OLlr(ogroups(name), D, Nt, Nx, 3, 0.0, E1, hw1, a0, seed, start=start, K=K, NRM=NRM, Ewin2=E2,
                       hw2=hw2, a20=a20, links=links)


@pytest.mark.parametrize("case", CASES2D, ids=lambda c: f"{c[0]}-{c[2]}x{c[3]}-{c[11]}")
@pytest.mark.parametrize("kernel", ["table", "class", "class-all-exact", "field", "field-all-exact"])
def test_llr2d_matches_oracle(dsa, dgroups, oracle, ogroups, case, kernel):
    say("case llr2d vs oracle", case, kernel)
    o = oracle2d(oracle, ogroups, case)
    with dsa.dsa.LlrEngine(4, 2) as eng:
        if kernel.endswith("-all-exact"):
            eng.set_kernel(kernel[: -len("-all-exact")], 1e300)
        else:
            eng.set_kernel(kernel)
        s = eng.add(dgroups(case[0]), params2d(dsa, case))
        for sweep in range(30):
            eng.step(1)
            o.step(1)
            same(eng.status(s), o.status())
            assert eng.status(s)["a2"] == o.status()["a2"]
            assert np.array_equal(eng.links(s), o.a_links), f"links differ after sweep {sweep + 1}"


def test_llr2d_from_seeded_cells_matches_oracle(dsa, dgroups, oracle, ogroups):
    """Starting from llr2d's own seeded cell configurations (the solve path)."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    from dym_symanzik_action_amd import cfgio

    cells = cfgio.parse_manifest([str(l) for l in f["manifest"]])[:3]
    say("case llr2d from seeded cells", len(cells))
    with dsa.dsa.LlrEngine(4, 4) as eng:
        slots, orc = [], []
        for k, c in enumerate(cells):
            links = f[os.path.basename(c["path"])[:-4]]
            case = ("mys1080-v4", 4, 3, 3, c["E1"], c["hw1"], c["E2"], c["hw2"], 1.0, 1.0, 300 + k, "hot", 5, 3)
            slots.append(eng.add(dgroups("mys1080-v4"), params2d(dsa, case), links=links))
            orc.append(oracle2d(oracle, ogroups, case, links=links))
        for _ in range(4):
            eng.step(5)
            for s, o in zip(slots, orc):
                o.step(5)
                same(eng.status(s), o.status())
                assert eng.status(s)["a2"] == o.status()["a2"]
                assert np.array_equal(eng.links(s), o.a_links)


def test_llr2d_statistics_match_reference_llr2d(dsa, dgroups):
    """llr2d's seed cells on 3^4 (tests/golden/llr2d_seed_3x3.npz), each solved
    with K = 20, NRM = 30 from a0 = a20 = 1.0: the reference's a_n over eight
    seeds and 16 device chains per cell agree for both tilts: Welch's t-test
    on the replicas, 2 x (number of cells) comparisons at the family-wise
    level of one two-sided 2-sigma test (Bonferroni)."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    from dym_symanzik_action_amd import cfgio

    cells = cfgio.parse_manifest([str(l) for l in f["manifest"]])
    ref1, ref2 = f["stats_a1"], f["stats_a2"]  # [seed, cell]
    nrep = 16
    say("case llr2d statistics", len(cells), "cells x", nrep)
    with dsa.dsa.LlrEngine(4, len(cells) * nrep) as eng:
        slots = []
        for k, c in enumerate(cells):
            links = f[os.path.basename(c["path"])[:-4]]
            for r in range(nrep):
                case = ("mys1080-v4", 4, 3, 3, c["E1"], c["hw1"], c["E2"], c["hw2"], 1.0, 1.0, 5000 + 100 * k + r,
                        "hot", 20, 30)
                slots.append(eng.add(dgroups("mys1080-v4"), params2d(dsa, case), links=links))
        for _ in range(400):
            eng.step(5)
            if all(eng.status(s)["phase"] == 2 for s in slots):
                break
        st = [eng.status(s) for s in slots]
    assert all(s["phase"] == 2 and s["reached"] == 1 and s["iterations"] == 30 for s in st)
    a1 = np.array([s["a"] for s in st]).reshape(len(cells), nrep)
    a2 = np.array([s["a2"] for s in st]).reshape(len(cells), nrep)
    from stats_util import p_limit, welch_p_samples

    worst = 1.0
    for k in range(len(cells)):
        for mine, ref in ((a1[k], ref1[:, k]), (a2[k], ref2[:, k])):
            p = welch_p_samples(mine, ref)
            say(f"  cell {k}: device {mine.mean():.5f}+-{mine.std(ddof=1):.5f}  "
                f"reference {ref.mean():.5f}+-{ref.std(ddof=1):.5f}  p={p:.4f}")
            worst = min(worst, p)
    assert worst > p_limit(2 * len(cells))
