"""Writes the golden fixtures of tests/golden/ from the COMPILED REFERENCE.

Run in the build container (needs /root/reference and `make -f oracle/ref.mk`):
    python tests/golden/make_golden.py
It runs oracle/_ref/dym-mod-metro (the reference driver built from its own
sources) and oracle/_ref/ref_harness (the reference's own get_staples,
get_rect_loops, update, plaquette, polyakov, wilson, getpoly called directly)
and stores their inputs and outputs as small .npz / .txt files.  Nothing here
runs where the GPU tests run: the tests only read the stored files.
"""
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("DSA_REFERENCE", "/root/reference")
REFBIN = os.path.join(REPO, "oracle", "_ref")


def run(cmd, **kw):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    kw.setdefault("cwd", tempfile.gettempdir())
    return subprocess.run(cmd, check=True, capture_output=True, text=True, env=env, **kw).stdout


def gmes_lines(text):
    return [l for l in text.splitlines() if l.startswith(("GMES:", "RMES:", "ACC:"))]


def loops_fixture(name, group, D, Nt, Nx, seed):
    P = int(open(os.path.join(REF, "groups", group)).read().split(None, 1)[0])
    V = Nt * Nx ** (D - 1)
    rng = np.random.default_rng(seed)
    cfg = rng.integers(0, P, size=V * D, dtype=np.int32)
    with tempfile.TemporaryDirectory() as td:
        cf = os.path.join(td, "cfg.i32")
        cfg.tofile(cf)
        run([os.path.join(REFBIN, "ref_harness"), "loops", os.path.join(REF, "groups", group), str(D), str(Nt),
             str(Nx), cf, td])
        ld = lambda f, t: np.fromfile(os.path.join(td, f), dtype=t)
        ns, nr = 2 * (D - 1), 6 * (D - 1)
        np.savez_compressed(
            os.path.join(HERE, f"loops_{name}.npz"),
            group=group, D=D, Nt=Nt, Nx=Nx, cfg=cfg,
            staples=ld("staples.i32", np.int32).reshape(V * D, ns),
            rect=ld("rect.i32", np.int32).reshape(V * D, nr),
            r1=ld("r1.f64", np.float64), m1=ld("m1.f64", np.float64),
            plaq=ld("plaq.i32", np.int32), rect21=ld("rect21.i32", np.int32),
            poly=ld("poly.i32", np.int32), wil=ld("wil.i32", np.int32),
            means=ld("means.f64", np.float64),
        )


def traj_fixture(name, group, D, Nt, Nx, b0, b1, b2, seed, nsweeps, start="hot", every=1):
    V = Nt * Nx ** (D - 1)
    with tempfile.TemporaryDirectory() as td:
        run([os.path.join(REFBIN, "ref_harness"), "traj", os.path.join(REF, "groups", group), str(D), str(Nt),
             str(Nx), repr(b0), repr(b1), repr(b2), str(seed), str(nsweeps), start, td, str(every)])
        steps = sorted(int(f[6:-4]) for f in os.listdir(td) if f.startswith("links_"))
        links = np.stack([np.fromfile(os.path.join(td, f"links_{s}.i32"), dtype=np.int32) for s in steps])
        counters = np.fromfile(os.path.join(td, "counters.u64"), dtype=np.uint64).reshape(-1, 3)
    np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), group=group, D=D, Nt=Nt, Nx=Nx,
                        beta=np.array([b0, b1, b2]), seed=seed, start=start, steps=np.array(steps),
                        links=links, counters=counters)
    assert links.shape[1] == V * D


def cli_fixture(name, args, threads=1):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    with tempfile.TemporaryDirectory() as td:
        out = subprocess.run([os.path.join(REFBIN, "dym-mod-metro")] + args, check=True, capture_output=True,
                             text=True, env=env, cwd=td).stdout
    lines = [l for l in out.splitlines() if not l.startswith("TM_")]
    with open(os.path.join(HERE, f"cli_{name}.txt"), "w") as f:
        f.write("# " + " ".join(["dym-mod-metro"] + args) + f"  (OMP_NUM_THREADS={threads})\n")
        f.write("\n".join(lines) + "\n")


def replicas_fixture(name, group, head, seeds, tail, nx, nt):
    """Per-replica means of the GMES/RMES observables over independent seeds."""
    rows = []
    for sd in seeds:
        with tempfile.TemporaryDirectory() as td:
            out = run([os.path.join(REFBIN, "dym-mod-metro"), os.path.join(REF, "groups", group)] + head +
                      [str(sd)] + tail, cwd=td)
        g = [l.split() for l in out.splitlines() if l.startswith("GMES:")]
        r = [l.split() for l in out.splitlines() if l.startswith("RMES:")]
        arr = np.array([[float(v) for v in row[1:]] for row in g])
        rect = np.array([float(row[2]) for row in r])
        rows.append(np.concatenate([[arr[:, 3].mean(), rect.mean(), np.hypot(arr[:, 1], arr[:, 2]).mean(),
                                     arr[:, 1].mean()], arr[:, 4:].mean(0)]))
    np.savez_compressed(os.path.join(HERE, f"replicas_{name}.npz"), group=group, args=np.array(head + tail),
                        seeds=np.array(list(seeds)), nx=nx, nt=nt, means=np.array(rows),
                        columns=np.array(["plaq", "rect", "absP", "ReP"] +
                                         [f"W{k + 1}x{l + 1}" for k in range(nx) for l in range(nt)]))


def llr_fixture(name, args):
    """stdout of the reference llr driver (llr.cpp) at one thread."""
    with tempfile.TemporaryDirectory() as td:
        out = run([os.path.join(REFBIN, "llr"), os.path.join(REF, "groups", args[0])] + args[1:], cwd=td)
    with open(os.path.join(HERE, f"llr_{name}.txt"), "w") as f:
        f.write("# llr " + " ".join(args) + "  (OMP_NUM_THREADS=1)\n")
        f.write(out)


def kentucky_fixture(name, group, head, seed, K, N):
    """SU(3) configurations written by the reference's savecfg driver
    (dym-mod-metro-savecfg.cpp:55-86) after every K sweeps, N of them."""
    with tempfile.TemporaryDirectory() as td:
        run([os.path.join(REFBIN, "dym-mod-metro-savecfg"), os.path.join(REF, "groups", group)] + head +
            [str(seed), td + "/", str(K), str(N), "0"], cwd=td)
        files = sorted(f for f in os.listdir(td) if f.startswith("lat-"))
        data = {f"file{n}": np.frombuffer(open(os.path.join(td, f), "rb").read(), np.uint8) for n, f in enumerate(files)}
    np.savez_compressed(os.path.join(HERE, f"kentucky_{name}.npz"), group=group, args=np.array(head),
                        seed=seed, K=K, N=N, names=np.array(files), **data)


def llr2d_fixture(name, group, D, Nt, Nx, seed_args, solve_args):
    """llr2d's seed pass (drive-in of every cell, CELL: manifest, raw int32
    .cfg per cell: llr2d.cpp:262-271, 416-474) and a solve pass over the
    manifest (ANE2: lines, llr2d.cpp:476-512)."""
    with tempfile.TemporaryDirectory() as td:
        cfgdir = os.path.join(td, "cfg")
        os.makedirs(cfgdir)
        out = run([os.path.join(REFBIN, "llr2d"), "seed", os.path.join(REF, "groups", group), str(D), str(Nt), str(Nx)]
                  + seed_args + [cfgdir], cwd=td)
        cells = [l for l in out.splitlines() if l.startswith("CELL:")]
        man = os.path.join(td, "manifest.txt")
        open(man, "w").write("\n".join(cells) + "\n")
        sol = run([os.path.join(REFBIN, "llr2d"), "solve", os.path.join(REF, "groups", group), str(D), str(Nt), str(Nx),
                   man, "0", str(len(cells))] + solve_args, cwd=td)
        cfgs = {}
        for l in cells:
            path = l.split()[-1]
            cfgs[os.path.basename(path)[:-4]] = np.fromfile(path, np.int32)
        # the same cells solved again with eight seeds (K = 20, NRM = 30): the
        # spread of the reference's a_n per cell, for statistical checks
        a1s, a2s = [], []
        for sd in range(100, 108):
            sol2 = run([os.path.join(REFBIN, "llr2d"), "solve", os.path.join(REF, "groups", group), str(D), str(Nt),
                        str(Nx), man, "0", str(len(cells)), solve_args[0], "20", "30", str(sd)], cwd=td)
            rows = [l.split() for l in sol2.splitlines() if l.startswith("ANE2:")]
            a1s.append([float(r[3]) for r in rows])
            a2s.append([float(r[4]) for r in rows])
    rel = [l.replace(cfgdir + "/", "cfg/") for l in cells]
    np.savez_compressed(os.path.join(HERE, f"llr2d_{name}.npz"), group=group, D=D, Nt=Nt, Nx=Nx,
                        seed_args=np.array(seed_args), solve_args=np.array(solve_args), manifest=np.array(rel),
                        ane2=np.array([l for l in sol.splitlines() if l.startswith("ANE2:")]),
                        head=np.array([l for l in out.splitlines() if l.startswith("LLR2D")]),
                        stats_a1=np.array(a1s), stats_a2=np.array(a2s), stats_seeds=np.arange(100, 108), **cfgs)


def llr_ladder_fixture(name, args):
    """The reference llr's ladder window by window (oracle/_ref/ref_llr_harness
    ladder: main's set-up, then one repeat of run_pass down and up calling
    llr.cpp's own seed_window / rm_solve): RECT_MULT, every window's energies
    and a_n as exact doubles, the links at each pass's start and after each
    window.  args = group D Nt Nx axis bg Etop Ebot step hw a0 seed K NRM."""
    with tempfile.TemporaryDirectory() as td:
        run([os.path.join(REFBIN, "ref_llr_harness"), "ladder", os.path.join(REF, "groups", args[0])] + args[1:] +
            [td])
        ld = lambda f, t: np.fromfile(os.path.join(td, f), dtype=t)
        win = ld("windows.f64", np.float64).reshape(-1, 9)
        links = {}
        for f in sorted(os.listdir(td)):
            if f.endswith(".i32"):
                links[f[:-4]] = ld(f, np.int32)
        np.savez_compressed(os.path.join(HERE, f"llrladder_{name}.npz"), args=np.array(args),
                            rect_mult=ld("init.f64", np.float64)[0], windows=win, **links)


def main():
    """Writes every fixture, or only the groups named on the command line
    (loops, traj, cli, replicas, llr, llrladder, files)."""
    if not os.path.exists(os.path.join(REFBIN, "ref_harness")):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    only = set(sys.argv[1:])
    want = lambda k: not only or k in only
    g = "mys1080-v4"
    if want("loops"):
        loops_fixture("s1080_2x2", g, 4, 2, 2, 11)
        loops_fixture("s1080_4x4", g, 4, 4, 4, 12)
        loops_fixture("s1080_6x4", g, 4, 6, 4, 13)
        loops_fixture("s216_4x6", "S216", 4, 4, 6, 14)
        loops_fixture("mybi_3x3", "myBI", 4, 3, 3, 15)
        loops_fixture("s108_d3_6x4", "S108", 3, 6, 4, 16)
    if want("traj"):  # trajectories of the reference update() at one thread
        traj_fixture("c1", g, 4, 2, 2, 0.0, 1.5, 0.1, 1, 110, every=10)
        traj_fixture("s1080_4x4_b2", g, 4, 4, 4, 0.0, 3.0, -0.25, 5, 6)
        traj_fixture("s1080_4x4_wilson", g, 4, 4, 4, 0.0, 2.0, 0.0, 3, 6)
        traj_fixture("s216_4x6", "S216", 4, 4, 6, 0.0, 2.5, -0.2, 21, 3)
        traj_fixture("mybi_3x3", "myBI", 4, 3, 3, 0.0, 1.0, 0.5, 77, 4)
    if want("cli"):  # full CLI runs (config 1; long chains where the Metropolis chain mixes)
        cli_fixture("c1", [os.path.join(REF, "groups", g), "4", "2", "2", "0", "1.5", "0.1", "1", "10", "100", "100"])
        cli_fixture("stats_4x4", [os.path.join(REF, "groups", g), "4", "4", "4", "0", "1.2", "-0.1", "9", "2",
                                  "1000", "50"])
        cli_fixture("stats_6x6", [os.path.join(REF, "groups", g), "4", "6", "6", "0", "1.6", "-0.25", "19", "1",
                                  "600", "50"])
    if want("replicas"):  # config 1 is frozen (acceptance 0.002): ensembles of independent replicas
        replicas_fixture("c1", g, ["4", "2", "2", "0", "1.5", "0.1"], range(1000, 1064), ["10", "10", "100"], 2, 2)
    if want("llr"):
        # llr.cpp: ladders of windows on axis 1 and 2 (exact-output fixtures),
        # and one window solved 16 times (8 repeats x hot/cold)
        llr_fixture("axis1_3x3", [g, "4", "3", "3", "1", "-0.1", "-500", "-800", "150", "60", "-1.0", "5", "3", "2",
                                  "1"])
        llr_fixture("axis2_4x3", [g, "4", "4", "3", "2", "1.2", "-300", "-500", "100", "50", "0.5", "6", "2", "2",
                                  "1"])
        llr_fixture("s216_axis1", ["S216", "4", "3", "4", "1", "0.0", "-900", "-1000", "100", "60", "1.0", "8", "2",
                                   "3", "1"])
        llr_fixture("stats_4x4", [g, "4", "4", "4", "1", "-0.1", "-1500", "-1500", "100", "75", "1.5", "20", "20",
                                  "30", "8"])
    if want("llrladder"):  # llr's ladders window by window, exact (the replay's pins)
        llr_ladder_fixture("axis1_3x3", [g, "4", "3", "3", "1", "-0.1", "-500", "-800", "150", "60", "-1.0", "5", "3",
                                         "2"])
        llr_ladder_fixture("axis2_4x3", [g, "4", "4", "3", "2", "1.2", "-300", "-500", "100", "50", "0.5", "6", "2",
                                         "2"])
        llr_ladder_fixture("s216_axis1_3x4", ["S216", "4", "3", "4", "1", "0.0", "-900", "-1000", "100", "60", "1.0",
                                              "8", "2", "3"])
        # a negative seed: llr seeds its engines with the unsigned sum iseed + i
        llr_ladder_fixture("axis1_3x3_negseed", [g, "4", "3", "3", "1", "0.0", "-450", "-600", "150", "60", "-0.8",
                                                 "-7", "2", "2"])
        # extent-2 lattices: rectangles wrap, RECT_MULT != 6 (7.77 on 2^4 and 7.69
        # on 2^3 x 3); the second ladder's up pass stops at an unreached window
        llr_ladder_fixture("axis1_2x2", [g, "4", "2", "2", "1", "0.0", "-100", "-150", "50", "20", "0.5", "9", "2",
                                         "2"])
        llr_ladder_fixture("axis2_3x2", [g, "4", "3", "2", "2", "0.8", "-50", "-100", "50", "25", "0.5", "12", "2",
                                         "2"])
    if want("files"):  # configuration files: savecfg's SU(3) output on config 1, llr2d's cells
        kentucky_fixture("c1", g, ["4", "2", "2", "0", "1.5", "0.1"], 1, 3, 2)
        llr2d_fixture("seed_3x3", g, 4, 3, 3,
                      ["-400", "-600", "100", "40", "0", "0.8", "0", "60", "40", "1", "1.0", "5"],
                      ["1.0", "10", "10", "7"])
    meta = {
        "generator": "tests/golden/make_golden.py",
        "reference_build": "oracle/ref.mk (g++ -O3 -fopenmp -std=c++11, as the reference Makefile)",
        "gcc": subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
        "glibc": platform.libc_ver()[1],
        "libstdc++": "GCC 11 headers (/usr/include/c++/11)",
        "threads": 1,
    }
    with open(os.path.join(HERE, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
