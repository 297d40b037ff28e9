"""The llr and llr2d drop-ins (the reference's llr.cpp and llr2d.cpp command
lines): stdout line for line in the reference's formats, every window and
cell bit-exact against the oracle replaying the same ladder or ridge walk
(its Philox keys read from the DSA_TRACE lines), the seed pass's cells and
order against the reference's own manifest (tests/golden/llr2d_seed_3x3.npz),
and a_n in statistical agreement with the reference llr's
(tests/golden/llr_stats_4x4.txt)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, group_file

pytestmark = pytest.mark.gpu

BUILD = os.path.join(REPO, "dym_symanzik_action_amd", "build")
LLR = os.path.join(BUILD, "llr")
LLR2D = os.path.join(BUILD, "llr2d")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(exe, args, cwd=None, trace=True):
    env = dict(os.environ)
    if trace:
        env["DSA_TRACE"] = "1"
    r = subprocess.run([exe] + args, capture_output=True, text=True, cwd=cwd, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout, r.stderr


def tasks(stderr):
    """label -> {seed, start, init, reached, a, a2} from the TASK lines."""
    out = {}
    for line in stderr.splitlines():
        if not line.startswith("TASK: "):
            continue
        head, _, rest = line[6:].partition(" seed=")
        kv = dict(f.split("=") for f in ("seed=" + rest).split())
        out[head] = {"seed": int(kv["seed"]), "start": int(kv["start"]), "init": int(kv["init"]),
                     "reached": int(kv["reached"]), "a": float(kv["a"]), "a2": float(kv["a2"])}
    return out


def finish(o):
    for _ in range(2000):
        if o.status()["phase"] == 2:
            return o.status()
        o.step(20)
    raise AssertionError("oracle chain did not finish")


def mean_sd(xs):
    """llr.cpp:336-338 / llr2d.cpp:489-492, in the same order of operations."""
    n = len(xs)
    m = 0.0
    for x in xs:
        m += x
    m /= n
    v = 0.0
    for x in xs:
        v += (x - m) * (x - m)
    return m, (np.sqrt(v / (n - 1)) if n > 1 else 0.0)


# ------------------------------------------------------------------- llr

def test_llr_ladders_match_oracle_replay(oracle, ogroups):
    """The 3-window ladder of tests/golden/llr_axis1_3x3.txt (3^4, S1 windows
    -500..-800, hw 60, a0 -1): header lines as the reference prints them, and
    every window of both ladders replayed by the oracle from the previous
    window's configuration and tilt gives the same a_n bit for bit."""
    ref = open(os.path.join(GOLDEN, "llr_axis1_3x3.txt")).read().splitlines()
    args = ref[0].split()[2:-1]  # "# llr <args>  (OMP...)"
    say("case llr cli", args)
    out, err = run(LLR, [group_file(args[0])] + args[1:])
    lines = out.splitlines()
    strip = lambda l: l.replace(group_file(args[0]), args[0]).replace("/root/reference/groups/" + args[0], args[0])
    assert [strip(l) for l in lines[:6]] == [strip(l) for l in ref[1:7]]
    T = tasks(err)
    D, Nt, Nx, axis = (int(x) for x in args[1:5])
    bg, Etop, Ebot, step, hw, a0 = (float(x) for x in args[5:11])
    K, NRM = int(args[12]), int(args[13])
    M = int((Etop - Ebot) / step + 0.5) + 1
    an = {n: [] for n in range(M)}
    for d in ("down", "up"):
        links, a_prev = None, a0
        for s in range(M):
            n = s if d == "down" else M - 1 - s
            t = T[f"0 {d} {n}"]
            o = oracle.OLlr(ogroups(args[0]), D, Nt, Nx, axis, bg, Etop - n * step, hw, a_prev, t["seed"],
                            start="hot" if d == "down" else "cold", K=K, NRM=NRM, cap=20000, links=links)
            st = finish(o)
            assert (st["reached"], st["a"]) == (t["reached"], t["a"]), (d, n)
            if not st["reached"]:
                break
            an[n].append(st["a"])
            links, a_prev = o.a_links, st["a"]
    from dym_symanzik_action_amd import cfgio

    want = [cfgio.ane_line(Etop - n * step, *mean_sd(an[n]), len(an[n])) if an[n] else
            cfgio.ane_line(Etop - n * step, 0, 0, 0) for n in range(M)]
    assert [l for l in lines if l.startswith("ANE:")] == want
    assert lines[-1].startswith(f"max |incr - recompute| E(axis{axis}) : ")


def test_llr_statistics_match_reference_llr_cli():
    """tests/golden/llr_stats_4x4.txt's command line (4^4, one window, R = 8
    repeats of both ladders): the drop-in's a_n agrees with the reference's
    (Welch's t-test over the 16 ladders each, two-sided 2-sigma level)."""
    ref = open(os.path.join(GOLDEN, "llr_stats_4x4.txt")).read().splitlines()
    args = ref[0].split()[2:-1]
    say("case llr cli statistics", args)
    out, _ = run(LLR, [group_file(args[0])] + args[1:], trace=False)
    mine = [l for l in out.splitlines() if l.startswith("ANE:")][0].split()
    theirs = [l for l in ref if l.startswith("ANE:")][0].split()
    assert mine[1] == theirs[1] and mine[4] == theirs[4] == "16"
    m1, s1, m2, s2 = float(mine[2]), float(mine[3]), float(theirs[2]), float(theirs[3])
    from stats_util import p_limit, welch_p

    p = welch_p(m1, s1, 16, m2, s2, 16)
    say(f"  a_n device {m1} +- {s1}, reference {m2} +- {s2}: p = {p:.3f}")
    assert p > p_limit(1)
    assert sum(l.startswith("repeat ") for l in out.splitlines()) == 8


# ----------------------------------------------------------------- llr2d

def grid(args):
    E1top, E1bot, step1, hw1, c0, c1, c2, step2, hw2 = (float(x) for x in args[:9])
    nperp = int(args[9])
    N1 = int((E1top - E1bot) / step1 + 0.5) + 1
    E1 = lambda i: E1top - i * step1
    E2 = lambda i, j: c0 + c1 * E1(i) + c2 * E1(i) * E1(i) + (j - nperp) * step2
    return N1, nperp, hw1, hw2, E1, E2


def replay_walk(oracle, og, D, Nt, Nx, g, cold, label, T, a0, K, NRM, on_cell):
    """One ridge walk of llr2d.cpp:324-396 (NRM = 0: the seed pass,
    llr2d.cpp:452-472) on the oracle, with the drop-in's Philox keys."""
    N1, nperp, hw1, hw2, E1, E2 = g
    links = None
    for s in range(N1):
        i = N1 - 1 - s if cold else s
        t = T[f"{label} {i} {nperp}"]
        o = oracle.OLlr(og, D, Nt, Nx, 3, 0.0, E1(i), hw1, a0, t["seed"], start="cold" if cold else "hot",
                        K=K, NRM=NRM, cap=4000, Ewin2=E2(i, nperp), hw2=hw2, a20=a0, links=links)
        st = finish(o)
        assert (st["reached"], st["a"], st["a2"]) == (t["reached"], t["a"], t["a2"]), (label, i)
        links = o.a_links
        if not st["reached"]:
            continue
        on_cell(i, nperp, st, links)
        for d in (-1, 1):
            for tt in range(1, nperp + 1):
                j = nperp + d * tt
                t = T[f"{label} {i} {j}"]
                c = oracle.OLlr(og, D, Nt, Nx, 3, 0.0, E1(i), hw1, st["a"], t["seed"], K=K, NRM=NRM, cap=4000,
                                Ewin2=E2(i, j), hw2=hw2, a20=st["a2"], links=links)
                ct = finish(c)
                assert (ct["reached"], ct["a"], ct["a2"]) == (t["reached"], t["a"], t["a2"]), (label, i, j)
                if not ct["reached"]:
                    break
                on_cell(i, j, ct, c.a_links)


def test_llr2d_full_grid_matches_oracle_replay(oracle, ogroups):
    """A 3 x 3 cell grid on 3^4 (E1 -400..-600, ridge E2 = 0.8 E1, nperp 1,
    K 3, NRM 2): the ANE2 lines equal the oracle's replay of the cold and the
    hot ridge walk, fans included."""
    grp = "mys1080-v4"
    args = ["-400", "-600", "100", "40", "0", "0.8", "0", "60", "40", "1", "1.0", "5", "3", "2", "1"]
    say("case llr2d cli full", args)
    out, err = run(LLR2D, [group_file(grp), "4", "3", "3"] + args)
    T = tasks(err)
    g = grid(args)
    N1, nperp = g[0], g[1]
    cells = {}
    reach = {}
    for cold, label in ((True, "0 cold"), (False, "0 hot")):
        ridge = []

        def on_cell(i, j, st, links):
            cells.setdefault((i, j), []).append((st["a"], st["a2"]))
            if j == nperp:
                ridge.append(i)

        replay_walk(oracle, ogroups(grp), 4, 3, 3, g, cold, label, T, 1.0, 3, 2, on_cell)
        reach[cold] = len(ridge)
    from dym_symanzik_action_amd import cfgio

    want = []
    for i in range(N1):
        for j in range(2 * nperp + 1):
            if (i, j) not in cells:
                continue
            xs = cells[(i, j)]
            m1, s1 = mean_sd([x[0] for x in xs])
            m2, s2 = mean_sd([x[1] for x in xs])
            want.append(cfgio.ane2_line(g[4](i), g[5](i, j), m1, m2, s1, s2, len(xs)))
    lines = out.splitlines()
    assert [l for l in lines if l.startswith("ANE2:")] == want
    assert f"repeat 0: cold ridge {reach[True]}/{N1}, hot ridge {reach[False]}/{N1}" in lines
    assert f"reached {len(want)}/{N1 * (2 * nperp + 1)} cells" in lines
    assert lines[0].startswith("LLR2D(grp,D,Nt,Nx): ")


def test_llr2d_seed_pass_matches_reference_manifest_and_oracle(oracle, ogroups, tmp_path):
    """llr2d seed with the arguments of tests/golden/llr2d_seed_3x3.npz: the
    same cells in the same order as the reference's manifest (both reach all
    nine), each saved configuration equal to the oracle's drive-in replay."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    grp = str(f["group"])
    seed_args = [str(x) for x in f["seed_args"]]
    cfgdir = tmp_path / "cfg"
    cfgdir.mkdir()
    say("case llr2d cli seed", seed_args)
    out, err = run(LLR2D, ["seed", group_file(grp), "4", "3", "3"] + seed_args + [str(cfgdir)])
    mine = [l for l in out.splitlines() if l.startswith("CELL:")]
    ref = [str(l) for l in f["manifest"]]
    assert [l.rsplit(" ", 1)[0] for l in mine] == [l.rsplit(" ", 1)[0] for l in ref]
    assert [l.rsplit(" ", 1)[1] for l in mine] == [str(cfgdir / os.path.basename(l.rsplit(" ", 1)[1])) for l in ref]
    assert f"SEED: saved {len(mine)} cells to {cfgdir}" in err
    T = tasks(err)
    g = grid(seed_args)
    expect = {}

    def on_cell(i, j, st, links):
        expect.setdefault((i, j), links)  # the first walk to reach a cell saves it

    for cold, label in ((True, "0 cold"), (False, "0 hot")):
        replay_walk(oracle, ogroups(grp), 4, 3, 3, g, cold, label, T, float(seed_args[10]), 1, 0, on_cell)
    from dym_symanzik_action_amd import cfgio

    for c in cfgio.parse_manifest(mine):
        assert np.array_equal(np.fromfile(c["path"], np.int32), expect[(c["i"], c["j"])]), c


def test_llr2d_solve_matches_oracle_on_reference_cells(oracle, ogroups, tmp_path):
    """llr2d solve over the reference's own seeded cells and manifest
    (relative paths, as the reference's seed pass wrote them): the cells of
    the reference's ANE2 lines, in its order, each a_n bit-exact against the
    oracle from the same configuration."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    grp = str(f["group"])
    (tmp_path / "cfg").mkdir()
    man = [str(l) for l in f["manifest"]]
    for l in man:
        p = l.rsplit(" ", 1)[1]
        f[os.path.basename(p)[:-4]].tofile(tmp_path / p)
    (tmp_path / "manifest.txt").write_text("# seeded\n" + "\n".join(man) + "\n")
    a0, K, NRM, seed = (str(x) for x in f["solve_args"])
    lo, hi = 1, len(man)
    say("case llr2d cli solve", a0, K, NRM, seed)
    out, err = run(LLR2D, ["solve", group_file(grp), "4", "3", "3", "manifest.txt", str(lo), str(hi), a0, K, NRM, seed],
                   cwd=str(tmp_path))
    rows = [l for l in out.splitlines() if l.startswith("ANE2:")]
    ref = [str(l) for l in f["ane2"]][lo:hi]
    assert [r.split()[1:3] for r in rows] == [r.split()[1:3] for r in ref]
    assert f"SOLVE[{lo},{hi}): {hi - lo} cells" in err
    T = tasks(err)
    from dym_symanzik_action_amd import cfgio

    cells = cfgio.parse_manifest(man)[lo:hi]
    want = []
    for k, c in enumerate(cells):
        t = T[f"cell {k}"]
        o = oracle.OLlr(ogroups(grp), 4, 3, 3, 3, 0.0, c["E1"], c["hw1"], float(a0), t["seed"], K=int(K), NRM=int(NRM),
                        cap=4000, Ewin2=c["E2"], hw2=c["hw2"], a20=float(a0), links=f[os.path.basename(c["path"])[:-4]])
        st = finish(o)
        want.append(cfgio.ane2_line(c["E1"], c["E2"], st["a"], st["a2"]))
    assert rows == want


def test_llr_axis2_two_repeats_match_oracle_replay(oracle, ogroups):
    """Axis 2 (the rectangle action windowed, bg = beta1) on 4^4 with two
    repeats of both ladders (R = 2): each of the four ladders replayed by the
    oracle, and the ANE lines' means and deviations over the four."""
    args = ["4", "4", "4", "2", "1.0", "-1200", "-1400", "100", "80", "0.5", "9", "3", "2", "2"]
    say("case llr cli axis 2, R = 2", args)
    out, err = run(LLR, [group_file("S216")] + args)
    T = tasks(err)
    D, Nt, Nx, axis = 4, 4, 4, 2
    bg, Etop, Ebot, step, hw, a0 = 1.0, -1200.0, -1400.0, 100.0, 80.0, 0.5
    K, NRM, R = 3, 2, 2
    M = int((Etop - Ebot) / step + 0.5) + 1
    an = {n: [] for n in range(M)}
    for r in range(R):
        for d in ("down", "up"):
            links, a_prev = None, a0
            for s in range(M):
                n = s if d == "down" else M - 1 - s
                t = T[f"{r} {d} {n}"]
                o = oracle.OLlr(ogroups("S216"), D, Nt, Nx, axis, bg, Etop - n * step, hw, a_prev, t["seed"],
                                start="hot" if d == "down" else "cold", K=K, NRM=NRM, cap=20000, links=links)
                st = finish(o)
                assert (st["reached"], st["a"]) == (t["reached"], t["a"]), (r, d, n)
                if not st["reached"]:
                    break
                an[n].append(st["a"])
                links, a_prev = o.a_links, st["a"]
    lines = out.splitlines()
    assert sum(l.startswith("repeat ") for l in lines) == R
    got = {float(l.split()[1]): l.split() for l in lines if l.startswith("ANE:") and "unreached" not in l}
    for n in range(M):
        if not an[n]:
            continue
        row = got[Etop - n * step]
        m, sd = mean_sd(an[n])
        # the drop-in collects the ladders as they finish: the same values in
        # another order, so the mean may differ in the last bit of %9.5f only
        assert abs(float(row[2]) - m) <= 1.01e-5 and abs(float(row[3]) - sd) <= 1.01e-5 and int(row[4]) == len(an[n])


def test_llr2d_solve_skips_missing_cells(tmp_path):
    """A manifest line whose configuration file is missing is reported on
    stderr and skipped (llr2d.cpp:500-501); the other cells are solved."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    (tmp_path / "cfg").mkdir()
    man = [str(l) for l in f["manifest"]][:3]
    for l in man[1:]:
        p = l.rsplit(" ", 1)[1]
        f[os.path.basename(p)[:-4]].tofile(tmp_path / p)
    (tmp_path / "m.txt").write_text("\n".join(man) + "\n")
    out, err = run(LLR2D, ["solve", group_file("mys1080-v4"), "4", "3", "3", "m.txt", "0", "3", "1.0", "2", "2", "5"],
                   cwd=str(tmp_path), trace=False)
    assert "solve: missing cfg cfg/" in err and "SOLVE[0,3): 2 cells" in err
    assert [l.split()[1:3] for l in out.splitlines() if l.startswith("ANE2:")] == \
        [l.split()[4:6] for l in man[1:]]


@pytest.mark.parametrize("fixture", ["llr_axis1_3x3.txt", "llr_axis2_4x3.txt", "llr_s216_axis1.txt"])
def test_llr_replay_prints_reference_output(fixture):
    """`llr --replay` (the reference's serial chain replayed on the device)
    prints the reference binary's stdout line for line: the header lines,
    RECT_MULT as calibrated, the repeat lines, every ANE line and the drift
    line (tests/golden/llr_*.txt, the reference llr at one thread)."""
    ref = open(os.path.join(GOLDEN, fixture)).read().splitlines()
    args = ref[0].split()[2:-1]
    say("case llr --replay", fixture, args)
    out, _ = run(LLR, ["--replay", group_file(args[0])] + args[1:], trace=False)
    strip = lambda l: l.replace(group_file(args[0]), args[0]).replace("/root/reference/groups/" + args[0], args[0])
    assert [strip(l) for l in out.splitlines()] == [strip(l) for l in ref[1:]]


def test_llr_rect_mult_line_is_the_references_calibration():
    """On lattices with an extent of 2 the rectangles wrap and the reference's
    calibrated RECT_MULT differs from 6 (llr.cpp:155-171, 318-321): the
    default drop-in prints the reference's value and its warning for axis 2
    (the value of tests/golden/llrladder_axis2_3x2.npz, calibrated by the
    reference's own calibrate_rect_mult on its hot start)."""
    f = np.load(os.path.join(GOLDEN, "llrladder_axis2_3x2.npz"))
    args = [str(x) for x in f["args"]]
    say("case llr RECT_MULT line", args)
    proc = subprocess.run([LLR, group_file(args[0])] + args[1:4] + ["2", "0.8", "-50", "-50", "50", "25", "0.5",
                                                                      args[11], "1", "1"],
                          capture_output=True, text=True, timeout=300)
    lines = proc.stdout.splitlines()
    assert f"RECT_MULT calibrated: {float(f['rect_mult']):.6f}" in lines
    assert f"WARNING: RECT_MULT={float(f['rect_mult']):.3f} != 6 -- the 1x2 rectangles wrap on this lattice" in lines
