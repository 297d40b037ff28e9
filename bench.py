#!/usr/bin/env python3
"""Benchmark: Metropolis link-updates/s of one S(1080) chain at 32^3 x 64.

A step is one full sweep: every one of the 8,388,608 links gets its 24
staples built and NHIT = 20 Metropolis hits (dym-mod-metro.cpp:146-197), in 32
race-free colour passes (one kernel launch each).  Workload: hot start seed
11, beta = (0, 3.0, -0.25), Philox streams, S(1080) tables resident on the
device before the timed region.

Prints one JSON line.  Besides the contract fields it carries
  roofline      the sweep kernel against HBM (66 algorithmic bytes of link
                traffic per link-update), with the time per launch from a
                HIP event pair on the chain's stream around the timed sweeps;
                against the L2 request rate
                that its table gathers meet (L1->L2 read requests per launch
                from the committed counters vs one request per L2 channel
                per clock); and against a random-gather probe (table gathers
                per link-update: 608 for the table kernel; 85 + 25 per
                accepted hit + 98 per exact re-decision for the class
                kernel; the probe's gathers each hit a random table row,
                as the sweep's do);
  cpu_baseline  the reference binary (oracle/_ref/dym-mod-metro, built from
                the reference's sources) timed on this host's cores for two
                sweeps of the same workload, by its own TM_END: update line.
"""
import argparse
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

D, NT, NX = 4, 64, 32
V = NT * NX ** 3
LINKS = V * D
BETA = (0.0, 3.0, -0.25)
SEED = 11
NHIT = 20
BYTES_PER_LINK_UPDATE = 66      # 64 B of link reads + 2 B write per updated link
GATHERS_TABLE = 608     # 84 staple + 24 initial + 20*25 hit-loop table gathers
GATHERS_CLASS = 85      # the 24 staple-times-link products, built directly; per hit: LDS only
GATHERS_CLASS_ACCEPT = 25   # new link + 24 staple products moved by the accepted factor
GATHERS_CLASS_EXACT = 98    # b', b^-1 and 4 per staple for an exact re-decision
KERNEL_NAME = {"table": "k_sweep<4>", "class": "k_sweep_cls<4, true>", "class-narrow": "k_sweep_cls<4, false>",
               "matrix": "k_sweep_mat<4>"}
HBM_PEAK_GBS = 8000.0
L2_REQ_PEAK = 16 * 8 * 2.4e9   # L2 requests/s: 16 channels per XCD, 8 XCDs, 2.4 GHz
PMC_FILE = {"table": "k_sweep_32x64_pmc.txt", "class-narrow": "k_sweep_cls_narrow_32x64_pmc.txt"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_profile(form, key):
    """A per-launch figure (HBM bytes, L1->L2 read requests, L2 hit rate) from
    the committed rocprofv3 --pmc summary of the kernel form that ran (None
    where none is committed)."""
    if form not in PMC_FILE:
        return None
    try:
        for line in open(os.path.join(REPO, "profiles", PMC_FILE[form])):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(group_gz, cores):
    """Two sweeps of the same workload by the reference binary, on `cores` threads."""
    ref = os.path.join(REPO, "oracle", "_ref", "dym-mod-metro")
    if not os.path.exists(ref):
        return None
    td = tempfile.mkdtemp(prefix="dsa_cpu_")
    try:
        gtxt = os.path.join(td, "mys1080-v4")
        with gzip.open(group_gz, "rb") as fi, open(gtxt, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        nsw = 2
        cmd = [ref, gtxt, str(D), str(NT), str(NX), *map(repr, BETA), str(SEED), str(nsw), "1", "0"]
        p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, env=env, cwd=td)
        t = None
        deadline = time.time() + 300
        for line in p.stdout:
            if line.startswith("TM_END:") and "update" in line:
                t = float(line.split("done in")[1].split()[0])
                break
            if time.time() > deadline:
                break
        p.kill()
        p.wait()
        if t is None:
            return None
        return {"value": nsw * LINKS / t, "unit": "link-updates/s", "cores": cores, "kind": "reference",
                "sample": f"{nsw} sweeps of 32^3x64 S1080 hot seed {SEED} beta={BETA} ({nsw} x 8,388,608 "
                          "link-updates), "
                          f"reference dym-mod-metro built from its sources (-O3 -fopenmp), OMP_NUM_THREADS={cores}, "
                          f"timed by its own TM_END: update line ({t:.2f} s)"}
    finally:
        shutil.rmtree(td, ignore_errors=True)


def scan_chains():
    """The phase-diagram scan workload: 256 chains on a 16 x 16 grid of
    beta1 in [1, 4] x beta2 in [-0.5, 0.3], shapes round-robin over 8^4, 12^4,
    16^3 x 8, 8^3 x 16, seeds 100 + i, hot; chains 192..255 use S(216)."""
    shapes = [(4, 8, 8), (4, 12, 12), (4, 8, 16), (4, 16, 8)]  # (D, Nt, Nx)
    out = []
    for i in range(256):
        b1 = 1.0 + 3.0 * (i // 16) / 15.0
        b2 = -0.5 + 0.8 * (i % 16) / 15.0
        out.append(("S216" if i >= 192 else "mys1080-v4", shapes[i % 4], b1, b2, 100 + i))
    return out


def batched_vs_serial(dsa, sweeps):
    """Link-updates/s of the 256-chain scan swept as one batch, and one chain
    at a time (each chain's own launches), over the same number of sweeps."""
    groups = {n: dsa.load_group(dsa.group_path(n)) for n in ("mys1080-v4", "S216")}
    chains = scan_chains()
    links = sum(s[1] * s[2] ** 3 * s[0] for _, s, *_ in chains)
    b = dsa.Batch()
    for name, (D, Nt, Nx), b1, b2, seed in chains:
        b.add(groups[name], D, Nt, Nx, 0.0, b1, b2, seed)
    b.finalize()  # times 256/512/1024 links per block and keeps the fastest
    task_links, tune_ms = b.task_links()
    b.sweep(1)
    t0 = time.perf_counter()
    b.sweep(sweeps)
    tb = time.perf_counter() - t0
    b.close()
    objs = []
    for name, (D, Nt, Nx), b1, b2, seed in chains:
        c = dsa.Chain(groups[name], D, Nt, Nx, 0.0, b1, b2, seed)
        c.update(1)
        objs.append(c)
    t0 = time.perf_counter()
    for c in objs:
        c.update(sweeps)
    ts = time.perf_counter() - t0
    for c in objs:
        c.close()
    return {"workload": "256 chains, 16x16 (beta1, beta2) grid, shapes 8^4/12^4/16^3x8/8^3x16 round-robin, "
                        "192 S(1080) + 64 S(216), hot seeds 100+i", "sweeps": sweeps, "links": links,
            "batched_link_updates_per_s": sweeps * links / tb, "serial_link_updates_per_s": sweeps * links / ts,
            "batched_s": tb, "serial_s": ts, "speedup": ts / tb, "task_links": task_links,
            "tuning_ms_per_sweep": {"256": tune_ms[0], "512": tune_ms[1], "1024": tune_ms[2]}}


def serving_requests(dsa, nreq, seed=2024):
    """LLR window requests as a Poisson stream (50 requests/s).  Each asks for
    a_n of one plaquette-energy window of width 1 % of the S1 range (NPLAQ *
    4.5 for S(1080), NPLAQ = 6V), centred uniformly on S1/NPLAQ in [-1.9, -0.6]
    (the greedy drive-in from a hot start stalls in metastable states below
    about -2.1 on 8^4); lattice 8^4, 12^4 or 16^4; beta2 in {0, -0.1, -0.25};
    hot start; drive-in cap 4000 sweeps (llr2d.cpp's); Robbins-Monro with K =
    20 sweeps per iteration until |delta a| < 1e-4 or 50 iterations
    (llr.cpp:223-250)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1 / 50.0, nreq - 1))])
    params, shapes = [], []
    for r in range(nreq):
        L = int(rng.choice([8, 12, 16]))
        b2 = float(rng.choice([0.0, -0.1, -0.25]))
        e = float(rng.uniform(-1.9, -0.6))
        nplaq = 6 * L ** 4
        a0 = 1.2 + (-0.7 - e) * (1.8 / 1.6)  # rough beta at which <S1>/NPLAQ = e
        params.append(dsa.dsa.llr_params(4, L, L, 1, b2, e * nplaq, 0.005 * 4.5 * nplaq, a0, 1000 + r, K=20, NRM=50,
                                         tol=1e-4, cap=4000))
        shapes.append(L)
    return arrivals, params, shapes


def serving(dsa, nreq, slots):
    """p50/p99 latency (arrival to converged a_n) and windows/s of the LLR
    serving loop, and windows/s serving the same requests one at a time."""
    import numpy as np

    g = dsa.load_group(dsa.group_path("mys1080-v4"))
    arrivals, params, shapes = serving_requests(dsa, nreq)
    with dsa.dsa.LlrEngine(4, slots) as eng:
        t0 = time.perf_counter()
        admit, done, st = eng.serve([g] * nreq, params, arrivals, chunk=5)
        makespan = time.perf_counter() - t0
    log(f"bench: served {nreq} windows in {makespan:.1f} s")
    lat = done - arrivals
    t_serial = 0.0
    for r, p in enumerate(params):
        with dsa.dsa.LlrEngine(4, 1) as eng:
            t0 = time.perf_counter()
            eng.serve([g], [p], [0.0], chunk=5)
            t_serial += time.perf_counter() - t0
        log(f"bench: window {r + 1}/{nreq} alone, {t_serial:.1f} s so far")
    return {"workload": f"{nreq} LLR window requests, Poisson 50/s, lattices 8^4/12^4/16^4 "
                        f"({shapes.count(8)}/{shapes.count(12)}/{shapes.count(16)}), beta2 in {{0,-0.1,-0.25}}, "
                        "window 1% of the S1 range, K=20, NRM=50, tol 1e-4",
            "slots": slots, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "windows_per_s": nreq / makespan,
            "serial_windows_per_s": nreq / t_serial, "speedup": t_serial / makespan,
            "reached": int(sum(s["reached"] for s in st)),
            "mean_rm_iterations": float(np.mean([s["iterations"] for s in st]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scan-sweeps", type=int, default=10,
                    help="sweeps of the 256-chain scan timed batched and serially (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="LLR window requests for the serving measurement (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--kernel", choices=["auto", "table", "class", "class-narrow", "matrix"], default="auto",
                    help="sweep kernel (identical results): element values via L2 gathers, or class values in LDS")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain copy kernel of this size (scale for the FETCH/WRITE_SIZE counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "independent chains shard across GPUs with no communication, see the design notes", file=sys.stderr)
        sys.exit(2)

    import dym_symanzik_action_amd as dsa

    gpath = dsa.group_path("mys1080-v4")
    g = dsa.load_group(gpath)
    log("bench: creating the 32^3x64 chain")
    c = dsa.Chain(g, D, NT, NX, *BETA, SEED, start="hot", rng="philox", order="colour")
    c.set_kernel(args.kernel)
    form = c.kernel_form()
    npass, ncol = c.schedule_info()
    log(f"bench: {npass} passes per sweep, colours per direction {ncol}; warmup {args.warmup}")
    import ctypes as C

    c.update(args.warmup)
    hit0, acc0 = c.counters()
    ex0 = c.exact_decisions()
    log(f"bench: timing {args.steps} sweeps")
    # one HIP event pair on the chain's stream around the timed sweeps: the
    # roofline's time per launch is their event time over the pass launches
    # (gaps between launches included).  Events around every launch would
    # order a cache release between passes and slow the sweep by ~9 %.
    dsa.lib().dsa_chain_set_profiling(c._h, 2)
    t0 = time.perf_counter()
    c.update(args.steps)          # returns after the stream has drained
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms, nl = C.c_double(), C.c_uint64()
    dsa.lib().dsa_chain_kernel_time(c._h, C.byref(ms), C.byref(nl))
    dsa.lib().dsa_chain_set_profiling(c._h, 0)
    hit1, acc1 = c.counters()
    ex1 = c.exact_decisions()
    assert hit1 - hit0 == args.steps * LINKS * NHIT
    value = args.steps * LINKS / elapsed
    avg_launch_s = ms.value * 1e-3 / nl.value
    units = LINKS // npass
    achieved = BYTES_PER_LINK_UPDATE * units / avg_launch_s / 1e9
    traffic = read_profile(form, "traffic_bytes_per_launch")
    l2_req = read_profile(form, "l2_read_requests_per_launch")
    l2_hit = read_profile(form, "l2_hit_rate")
    if form == "table":
        gathers_per_unit = float(GATHERS_TABLE)
    elif form == "matrix":
        # the 24 staple-times-link products (85), the new link of an accepted
        # hit; an exact re-decision builds the products again (85) and
        # evaluates them (98)
        gathers_per_unit = GATHERS_CLASS + ((acc1 - acc0) + (GATHERS_CLASS + GATHERS_CLASS_EXACT) *
                                            (ex1 - ex0)) / (args.steps * LINKS)
    else:
        gathers_per_unit = GATHERS_CLASS + (GATHERS_CLASS_ACCEPT * (acc1 - acc0) +
                                            GATHERS_CLASS_EXACT * (ex1 - ex0)) / (args.steps * LINKS)
    gather_rate = gathers_per_unit * units / avg_launch_s
    gp = C.c_double()
    best = 0.0
    for blocks in (2048, 4096, 8192):
        if dsa.lib().dsa_probe_gather(g._h, blocks, 64, C.byref(gp)) == 0:
            best = max(best, gp.value)
    m = c.measure(0, 0)
    c.close()
    copy_ms = None
    if args.copy_probe_mb:
        cm = C.c_double()
        if dsa.lib().dsa_probe_copy(args.copy_probe_mb << 20, 4, C.byref(cm)) == 0:
            copy_ms = cm.value

    scan = None
    if args.scan_sweeps:
        log(f"bench: 256-chain scan, {args.scan_sweeps} sweeps batched and one chain at a time")
        scan = batched_vs_serial(dsa, args.scan_sweeps)

    serve = None
    if args.serve_requests:
        log(f"bench: LLR serving, {args.serve_requests} requests, {args.serve_slots} slots, then one at a time")
        serve = serving(dsa, args.serve_requests, args.serve_slots)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cores = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            cores = min(16, os.cpu_count() or 1)
        log(f"bench: CPU baseline, reference binary on {cores} threads")
        cpu = cpu_baseline(gpath, cores)

    out = {
        "metric": "Metropolis link-updates/s at 32^3x64 S1080; batched-vs-serial chain speedup",
        "value": value,
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16 group indices, f64 action sums",
        "data": "synthetic: hot start from Philox seed 11 (no dataset involved)",
        "config": {"workload": "single chain 32^3x64 (8,388,608 uint16 links), S(1080) table, hot seed 11, "
                               "beta=(0,3.0,-0.25), NHIT=20, Philox4x32-10, race-free colour order",
                   "lattice": "32^3x64", "group": "mys1080-v4", "links": LINKS, "passes_per_sweep": npass,
                   "parallelism": "single GPU"},
        "acceptance": (acc1 - acc0) / (hit1 - hit0),
        "plaquette_after": m[2],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAME[form], "avg_launch_ms": avg_launch_s * 1e3, "units_per_launch": units,
                     "algorithmic_bytes_per_unit": BYTES_PER_LINK_UPDATE,
                     "gather": {"achieved_per_s": gather_rate, "ceiling_per_s": best,
                                "frac": gather_rate / best if best else None,
                                "per_unit": gathers_per_unit,
                                # the reference's arithmetic needs 608 per link-update:
                                # the rate it would have to sustain for this throughput
                                "reference_equivalent_per_s": GATHERS_TABLE * value,
                                "reference_equivalent_frac": GATHERS_TABLE * value / best if best else None},
                     # the bound the staple-product gathers meet: L1->L2 read
                     # requests (one 128-byte line each) per launch from the
                     # committed counters, over the live launch time, against
                     # one request per L2 channel per clock (16 x 8 XCDs x 2.4 GHz)
                     "l2_requests": {"achieved_per_s": l2_req / avg_launch_s if l2_req else None,
                                     "peak_per_s": L2_REQ_PEAK,
                                     "frac": l2_req / avg_launch_s / L2_REQ_PEAK if l2_req else None,
                                     "per_launch": l2_req, "l2_hit_rate": l2_hit},
                     "exact_redecisions_per_hit": (ex1 - ex0) / (hit1 - hit0)},
        "cpu_baseline": cpu,
        "batched_vs_serial": scan,
        "llr_serving": serve,
    }
    if copy_ms is not None:
        out["copy_probe"] = {"bytes": args.copy_probe_mb << 20, "ms": copy_ms}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
