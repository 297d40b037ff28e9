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
                and, under "limiter", what bounds it: the L2 request rate
                its table gathers meet (L1->L2 read requests per launch from
                the committed rocprofv3 counters vs one request per L2
                channel per clock) and a random-gather probe (table gathers
                per link-update: 608 for the table kernel; 85 + 25 per
                accepted hit + 98 per exact re-decision for the class
                kernel; 25 per accepted hit + 98 per re-decision for the
                field kernel; the probe's gathers each hit a random table
                row, as the sweep's do).  The field kernel builds its
                staple products as GF(4) matrix products in registers, so
                its limiter is the VALU issue rate instead: VALU wave-
                instructions per launch from the committed counters over the
                live launch time, against one per CU per shader clock;
  beta_scan     the 32^3 x 64 lattice at beta1 in {2.0, 2.5, 3.0, 3.5} x
                beta2 in {0, -0.25}, cold and hot (seed 11): link-updates/s,
                acceptance and exact re-decisions per hit, two sweeps each;
  batched_vs_serial  the 256-chain phase-diagram scan run as 30 hipGraph-
                captured cycles of 10 sweeps + measure (300 sweeps), against
                the same chains swept one at a time;
  llr_serving   a Poisson stream of LLR window requests served with
                continuous batching under an HBM budget: p50/p99 latency to a
                converged a_n, windows/s, and a sample of the requests one
                at a time;
  cpu_baseline  the reference binary (oracle/_ref/dym-mod-metro, built from
                the reference's sources) timed on this host for three sweeps
                of the same workload by its own TM_END: update line, on the
                host threads this process may use (at most 16), with a
                one-thread run beside it.
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

from dym_symanzik_action_amd import workloads  # noqa: E402

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
               "matrix": "k_sweep_mat<4>", "field": "k_sweep_cls<4, false, false, true>"}
HBM_PEAK_GBS = 8000.0
L2_REQ_PEAK = 16 * 8 * 2.4e9   # L2 requests/s: 16 channels per XCD, 8 XCDs, 2.4 GHz
VALU_ISSUE_PER_CU = 1           # wave64 VALU instructions per CU per clock (4 SIMDs x 16 lanes)
CUS = 256
PMC_FILE = {"table": "k_sweep_32x64_pmc.txt", "class-narrow": "k_sweep_cls_narrow_32x64_pmc.txt",
            "field": "k_sweep_cls_field_32x64_pmc.txt"}


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


def host_cpu():
    """CPU model, physical cores, logical CPUs and the CPUs this process may use."""
    model, phys, logical = None, set(), 0
    cur = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name" and model is None:
                    model = v
            elif cur:
                logical += 1
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
        if cur:
            logical += 1
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "physical_cores": len(phys) or None, "logical_cpus": logical or None,
            "usable_cpus": usable}


def reference_sweeps(gtxt, td, shape, nsw, threads, timeout_s=600):
    """Seconds the reference binary spends in `nsw` sweeps of the benchmark's
    couplings and seed on `shape` = (Nt, Nx), by its own TM_END: update line
    (the process is stopped there: its Nx x Nt Wilson grid measurement would
    take hours at 32^3 x 64)."""
    ref = os.path.join(REPO, "oracle", "_ref", "dym-mod-metro")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    nt, nx = shape
    cmd = [ref, gtxt, str(D), str(nt), str(nx), *map(repr, BETA), str(SEED), str(nsw), "1", "0"]
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, env=env, cwd=td)
    t = None
    deadline = time.time() + timeout_s
    for line in p.stdout:
        if line.startswith("TM_END:") and "update" in line:
            t = float(line.split("done in")[1].split()[0])
            break
        if time.time() > deadline:
            break
    p.kill()
    p.wait()
    return t


def cpu_baseline(group_gz, threads, nsw=3, nsw1=3, shape1=(32, 16)):
    """The reference binary on this host: `nsw` sweeps of the benchmark
    workload on `threads` OpenMP threads, and `nsw1` sweeps on one thread of
    the same couplings on the 16^3 x 32 lattice (one thread takes ~40 s per
    32^3 x 64 sweep)."""
    ref = os.path.join(REPO, "oracle", "_ref", "dym-mod-metro")
    if not os.path.exists(ref):
        return None
    td = tempfile.mkdtemp(prefix="dsa_cpu_")
    try:
        gtxt = os.path.join(td, "mys1080-v4")
        with gzip.open(group_gz, "rb") as fi, open(gtxt, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        t = reference_sweeps(gtxt, td, (NT, NX), nsw, threads)
        if t is None:
            return None
        links1 = D * shape1[0] * shape1[1] ** (D - 1)
        t1 = reference_sweeps(gtxt, td, shape1, nsw1, 1)
        cpu = host_cpu()
        return {"value": nsw * LINKS / t, "unit": "link-updates/s", "cores": threads, "kind": "reference",
                "sample": f"{nsw} sweeps of 32^3x64 S1080 hot seed {SEED} beta={BETA} ({nsw} x 8,388,608 "
                          f"link-updates), reference dym-mod-metro built from its sources (-O3 -fopenmp), "
                          f"OMP_NUM_THREADS={threads}, timed by its own TM_END: update line ({t:.2f} s)",
                "host": cpu,
                "one_thread": {"value": nsw1 * links1 / t1 if t1 else None, "unit": "link-updates/s", "cores": 1,
                               "sample": f"{nsw1} sweeps of {shape1[1]}^3x{shape1[0]} S1080, same couplings and "
                                         f"seed, OMP_NUM_THREADS=1 ({t1:.2f} s)" if t1 else "timed out"}}
    finally:
        shutil.rmtree(td, ignore_errors=True)


def beta_scan(dsa, g, sweeps=2):
    """The 32^3 x 64 lattice over the beta grid of the large-shape workload:
    beta1 in {2.0, 2.5, 3.0, 3.5} x beta2 in {0, -0.25}, cold and hot (seed
    11); per point, link-updates/s over `sweeps` sweeps after one untimed
    sweep, acceptance and exact re-decisions per hit."""
    rows = []
    for start in ("cold", "hot"):
        for b2 in (0.0, -0.25):
            for b1 in (2.0, 2.5, 3.0, 3.5):
                with dsa.Chain(g, D, NT, NX, 0.0, b1, b2, SEED, start=start) as c:
                    c.update(1)
                    h0, a0 = c.counters()
                    e0 = c.exact_decisions()
                    t0 = time.perf_counter()
                    c.update(sweeps)
                    dt = time.perf_counter() - t0
                    h1, a1 = c.counters()
                    e1 = c.exact_decisions()
                rows.append({"start": start, "beta1": b1, "beta2": b2, "link_updates_per_s": sweeps * LINKS / dt,
                             "acceptance": (a1 - a0) / (h1 - h0), "exact_redecisions_per_hit": (e1 - e0) / (h1 - h0)})
                log(f"bench: beta scan {start} ({b1}, {b2}): {rows[-1]['link_updates_per_s']:.3e} link-updates/s, "
                    f"acceptance {rows[-1]['acceptance']:.4f}")
    lu = [r["link_updates_per_s"] for r in rows]
    return {"workload": f"32^3x64 S(1080), beta1 in {{2.0,2.5,3.0,3.5}} x beta2 in {{0,-0.25}}, cold and hot seed "
                        f"{SEED}; {sweeps} timed sweeps per point after one untimed", "points": rows,
            "min_link_updates_per_s": min(lu), "max_link_updates_per_s": max(lu)}


def batched_vs_serial(dsa, cycles, sweeps_per_cycle, serial_sweeps):
    """Link-updates/s of the 256-chain scan as one batch -- `cycles`
    hipGraph-captured cycles of `sweeps_per_cycle` sweeps + measure -- and
    one chain at a time (each chain's own sweep launches, `serial_sweeps`
    sweeps per chain)."""
    groups = {n: dsa.load_group(dsa.group_path(n)) for n in ("mys1080-v4", "S216")}
    chains = workloads.scan_chains()
    links = workloads.scan_links(chains)
    b = dsa.Batch()
    for name, (d, nt, nx), b1, b2, seed in chains:
        b.add(groups[name], d, nt, nx, 0.0, b1, b2, seed)
    b.finalize()  # times 256/512/1024 links per block and keeps the fastest
    task_links, tune_ms = b.task_links()
    b.run_cycles(1, sweeps_per_cycle)  # capture the graph, one untimed cycle
    t0 = time.perf_counter()
    meas = b.run_cycles(cycles, sweeps_per_cycle)
    tb = time.perf_counter() - t0
    b.close()
    objs = []
    for name, (d, nt, nx), b1, b2, seed in chains:
        c = dsa.Chain(groups[name], d, nt, nx, 0.0, b1, b2, seed)
        c.update(1)
        objs.append(c)
    t0 = time.perf_counter()
    for c in objs:
        c.update(serial_sweeps)
    ts = time.perf_counter() - t0
    for c in objs:
        c.close()
    nsw = cycles * sweeps_per_cycle
    rb, rs = nsw * links / tb, serial_sweeps * links / ts
    return {"workload": "256 chains, 16x16 (beta1, beta2) grid, shapes 8^4/12^4/16^3x8/8^3x16 round-robin, "
                        "192 S(1080) + 64 S(216), hot seeds 100+i",
            "batched": f"{cycles} hipGraph replays of ({sweeps_per_cycle} sweeps + measure) = {nsw} sweeps per chain",
            "serial": f"each chain alone, {serial_sweeps} sweeps per chain, one chain after another",
            "links": links, "batched_link_updates_per_s": rb, "serial_link_updates_per_s": rs,
            "batched_s": tb, "serial_s": ts, "speedup": rb / rs, "task_links": task_links,
            "measured_plaquette_range": [float(meas[-1, :, 2].min()), float(meas[-1, :, 2].max())],
            "tuning_ms_per_sweep": {"256": tune_ms[0], "512": tune_ms[1], "1024": tune_ms[2]}}


def serving(dsa, nreq, slots, budget, serial_every):
    """p50/p99 latency (arrival to converged a_n) and windows/s of the LLR
    serving loop under an HBM budget, and windows/s of every
    `serial_every`-th request served alone, one after another."""
    import numpy as np

    g = dsa.load_group(dsa.group_path("mys1080-v4"))
    arrivals, params, shapes, centres = workloads.llr_requests(nreq)
    with dsa.dsa.LlrEngine(4, slots, hbm_budget=budget) as eng:
        t0 = time.perf_counter()
        admit, done, st = eng.serve([g] * nreq, params, arrivals, chunk=5)
        makespan = time.perf_counter() - t0
    log(f"bench: served {nreq} windows in {makespan:.1f} s")
    lat = done - arrivals
    fps = [dsa.dsa.llr_footprint(p) for p in params]
    peak = max(sum(f for a, d, f in zip(admit, done, fps) if a <= t < d) for t in admit)
    sample = list(range(0, nreq, serial_every))
    t_serial = 0.0
    for k, r in enumerate(sample):
        with dsa.dsa.LlrEngine(4, 1) as eng:
            t0 = time.perf_counter()
            eng.serve([g], [params[r]], [0.0], chunk=5)
            t_serial += time.perf_counter() - t0
        log(f"bench: window {k + 1}/{len(sample)} alone, {t_serial:.1f} s so far")
    served_rate = nreq / makespan
    serial_rate = len(sample) / t_serial
    return {"workload": f"{nreq} LLR window requests, Poisson 50/s, lattices 8^4/12^4/16^4 "
                        f"({shapes.count(8)}/{shapes.count(12)}/{shapes.count(16)}), beta2 in {{0,-0.1,-0.25}}, "
                        "window centres uniform over the S1 range [-3, 1.5] x NPLAQ, width 1% of the range, "
                        "cold start below -1.5 per plaquette, K=20, NRM=50, tol 1e-4, drive-in cap 200",
            "slots": slots, "hbm_budget_bytes": budget, "peak_admitted_bytes": peak,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "windows_per_s": served_rate, "serial_windows_per_s": serial_rate,
            "serial_sample": f"every {serial_every}th request alone ({len(sample)} requests)",
            "speedup": served_rate / serial_rate,
            "reached": int(sum(s["reached"] for s in st)), "rejected": int(sum(s["phase"] == 4 for s in st)),
            "mean_rm_iterations": float(np.mean([s["iterations"] for s in st]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="OpenMP threads of the reference CPU baseline (capped by the CPUs this process may use)")
    ap.add_argument("--beta-scan-sweeps", type=int, default=2,
                    help="timed sweeps per point of the 32^3x64 beta grid (0: skip)")
    ap.add_argument("--scan-cycles", type=int, default=30,
                    help="hipGraph cycles (10 sweeps + measure) of the 256-chain scan (0: skip)")
    ap.add_argument("--scan-serial-sweeps", type=int, default=10,
                    help="sweeps per chain of the one-chain-at-a-time comparison")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="LLR window requests for the serving measurement (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-budget-mb", type=int, default=128,
                    help="HBM budget of the serving engine in MiB (admission by footprint)")
    ap.add_argument("--serve-serial-every", type=int, default=5,
                    help="time every n-th request alone for the one-at-a-time rate")
    ap.add_argument("--kernel", choices=["auto", "table", "class", "class-narrow", "matrix", "field"], default="auto",
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
    valu = read_profile(form, "valu_insts_per_launch")
    lds_active = read_profile(form, "lds_active_cycles_per_launch")
    clock = read_profile(form, "shader_clock_ghz")
    if form == "table":
        gathers_per_unit = float(GATHERS_TABLE)
    elif form == "matrix":
        # the 24 staple-times-link products (85), the new link of an accepted
        # hit; an exact re-decision builds the products again (85) and
        # evaluates them (98)
        gathers_per_unit = GATHERS_CLASS + ((acc1 - acc0) + (GATHERS_CLASS + GATHERS_CLASS_EXACT) *
                                            (ex1 - ex0)) / (args.steps * LINKS)
    elif form == "field":
        # the staple-times-link products are GF(4) matrix products in
        # registers: table gathers only for accepted hits and re-decisions
        gathers_per_unit = (GATHERS_CLASS_ACCEPT * (acc1 - acc0) +
                            GATHERS_CLASS_EXACT * (ex1 - ex0)) / (args.steps * LINKS)
    else:
        gathers_per_unit = GATHERS_CLASS + (GATHERS_CLASS_ACCEPT * (acc1 - acc0) +
                                            GATHERS_CLASS_EXACT * (ex1 - ex0)) / (args.steps * LINKS)
    gather_rate = gathers_per_unit * units / avg_launch_s
    valu_issue = None
    if valu and clock:
        # VALU wave-instructions per second against one per CU per clock at
        # the shader clock the counters saw; LDS busy share of the CU cycles
        peak = CUS * VALU_ISSUE_PER_CU * clock * 1e9
        valu_issue = {"achieved_per_s": valu / avg_launch_s, "peak_per_s": peak,
                      "frac": valu / avg_launch_s / peak, "per_launch": valu,
                      "lds_busy_frac": lds_active / (CUS * clock * 1e9 * avg_launch_s) if lds_active else None}
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

    bscan = None
    if args.beta_scan_sweeps:
        log(f"bench: 32^3x64 beta grid, {args.beta_scan_sweeps} sweeps per point")
        bscan = beta_scan(dsa, g, args.beta_scan_sweeps)

    scan = None
    if args.scan_cycles:
        log(f"bench: 256-chain scan, {args.scan_cycles} graph cycles of 10 sweeps + measure, then one chain at a "
            "time")
        scan = batched_vs_serial(dsa, args.scan_cycles, 10, args.scan_serial_sweeps)

    serve = None
    if args.serve_requests:
        log(f"bench: LLR serving, {args.serve_requests} requests, {args.serve_slots} slots, budget "
            f"{args.serve_budget_mb} MiB, then a sample one at a time")
        serve = serving(dsa, args.serve_requests, args.serve_slots, args.serve_budget_mb << 20,
                        args.serve_serial_every)

    cpu = None
    if not args.no_cpu_baseline:
        threads = max(1, min(args.cpu_threads, host_cpu()["usable_cpus"]))
        log(f"bench: CPU baseline, reference binary on {threads} threads, then one thread")
        cpu = cpu_baseline(gpath, threads)

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
                     # what actually bounds the sweep: for the gather forms its
                     # table gathers, served by L2.  L1->L2 read requests (one
                     # 128-byte line each) per launch from the committed
                     # rocprofv3 counters, over the live launch time, against
                     # one request per L2 channel per clock (16 channels x 8
                     # XCDs x 2.4 GHz); and the gather rate against the
                     # random-gather probe's.  For the field form the VALU
                     # issue rate (the GF(4) products)
                     "limiter": {"kind": "valu_issue" if form == "field" else "l2_requests",
                                 "frac": ((valu_issue or {}).get("frac") if form == "field" else
                                          l2_req / avg_launch_s / L2_REQ_PEAK if l2_req else None),
                                 "valu_issue": valu_issue,
                                 "l2_requests": {"achieved_per_s": l2_req / avg_launch_s if l2_req else None,
                                                 "peak_per_s": L2_REQ_PEAK, "per_launch": l2_req,
                                                 "l2_hit_rate": l2_hit},
                                 "gather": {"achieved_per_s": gather_rate, "ceiling_per_s": best,
                                            "frac": gather_rate / best if best else None,
                                            "per_unit": gathers_per_unit,
                                            # the reference's arithmetic needs 608 per link-update:
                                            # the rate it would have to sustain for this throughput
                                            "reference_equivalent_per_s": GATHERS_TABLE * value,
                                            "reference_equivalent_frac": GATHERS_TABLE * value / best if best
                                            else None}},
                     "exact_redecisions_per_hit": (ex1 - ex0) / (hit1 - hit0)},
        "cpu_baseline": cpu,
        "beta_scan": bscan,
        "batched_vs_serial": scan,
        "llr_serving": serve,
    }
    if copy_ms is not None:
        out["copy_probe"] = {"bytes": args.copy_probe_mb << 20, "ms": copy_ms}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
