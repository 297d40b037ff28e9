"""The multi-chain workloads and the single-chain benchmark workload, checked
at their own sizes.

* The 256-chain phase-diagram scan (dym_symanzik_action_amd/workloads.py,
  the replacement of evermore.sh:20 / jobs2run's one process per point):
  one chain of every (shape, group) pair equals the oracle after a sweep,
  every chain equals the same chain run alone, and hipGraph-captured
  10-sweep + measure cycles equal the eager sequence.
* LLR serving under an HBM budget (the replacement of
  scripts/llr2d_parallel.sh:28-37): admission waits on footprint, the
  admitted footprints never exceed the budget, every served result equals
  the oracle's, and a request larger than the whole budget is rejected
  while the others are served.
* The benchmark workload itself, 32^3 x 64 S(1080): the default class
  kernel's trajectory equals the table kernel's (the reference's arithmetic,
  pinned to the oracle) bit for bit, counters included, hot and cold.
"""
import sys

import numpy as np
import pytest

from dym_symanzik_action_amd import workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scan_batch(dsa, dgroups):
    b = dsa.Batch()
    for name, (D, Nt, Nx), b1, b2, seed in workloads.scan_chains():
        b.add(dgroups(name), D, Nt, Nx, 0.0, b1, b2, seed)
    return b.finalize()


def test_scan_batch_matches_oracle_and_lone_chains(dsa, dgroups, oracle, ogroups):
    """All 256 chains of the scan in one batch (block size tuned), one sweep:
    the first chain of each (shape, group) pair -- 8 chains, 4 shapes x
    S(1080)/S(216) -- equals the oracle's links and counters, and every
    chain equals the same chain swept alone."""
    chains = workloads.scan_chains()
    say("case scan batch: 256 chains, 1 sweep")
    with scan_batch(dsa, dgroups) as b:
        assert b.info()[0] == 256
        b.sweep(1)
        hits, accs = b.counters()
        firsts = {}
        for c, (name, shape, *_rest) in enumerate(chains):
            firsts.setdefault((name, shape), c)
        assert len(firsts) == 8
        for (name, (D, Nt, Nx)), c in sorted(firsts.items(), key=lambda kv: kv[1]):
            _, _, b1, b2, seed = chains[c]
            say(f"  chain {c}: {name} {Nt}x{Nx}^3 beta=({b1:.3f}, {b2:.3f}) vs oracle")
            o = oracle.OChain(ogroups(name), D, Nt, Nx, 0.0, b1, b2, seed, start="hot", rng="philox",
                              order="colour")
            o.sweep(1)
            assert np.array_equal(b.links(c), o.a), f"chain {c} differs from the oracle"
            assert (int(hits[c]), int(accs[c])) == o.counters(), c
        say("  all 256 chains vs lone chains")
        for c, (name, (D, Nt, Nx), b1, b2, seed) in enumerate(chains):
            with dsa.Chain(dgroups(name), D, Nt, Nx, 0.0, b1, b2, seed) as ch:
                ch.update(1)
                assert np.array_equal(b.links(c), ch.a), f"chain {c} differs from the lone chain"
                assert (int(hits[c]), int(accs[c])) == ch.counters(), c


def test_scan_graph_cycles_equal_eager(dsa, dgroups):
    """The scan's schedule: 10-sweep + measure cycles captured in a hipGraph
    and replayed (three replays, with host reads between them) equal the
    same sweeps and measurements issued eagerly, for every chain."""
    say("case scan cycles: 3 x (10 sweeps + measure), graph vs eager")
    with scan_batch(dsa, dgroups) as eager, scan_batch(dsa, dgroups) as graph:
        want = []
        for _ in range(3):
            eager.sweep(10)
            want.append(eager.measure())
        got = [graph.run_cycles(1, 10)[0]]
        got.append(graph.run_cycles(2, 10))
        got = np.concatenate([got[0][None], got[1]])
        assert np.array_equal(got, np.array(want))
        for c in range(256):
            assert np.array_equal(graph.links(c), eager.links(c)), c
        assert all(np.array_equal(x, y) for x, y in zip(graph.counters(), eager.counters()))


# served requests: (L, beta2, centre per plaquette, start, seed); small K and
# NRM keep the oracle's replay short, the lattices are the serving workload's
SERVED = [
    (12, -0.1, -0.30, "hot", 41),
    (8, 0.0, -0.35, "hot", 42),
    (8, -0.25, -2.40, "cold", 43),
    (8, -0.1, -0.50, "hot", 44),
    (8, 0.0, -1.00, "hot", 45),
]


def served_params(dsa, L, b2, e, start, seed, K=4, NRM=3, cap=200):
    nplaq = 6 * L ** 4
    return dsa.dsa.llr_params(4, L, L, 1, b2, e * nplaq, 0.0225 * nplaq, 1.2 + (-0.7 - e) * 1.125, seed,
                              start=start, K=K, NRM=NRM, cap=cap)


def oracle_served(oracle, ogroups, L, b2, e, start, seed, K=4, NRM=3, cap=200):
    nplaq = 6 * L ** 4
    o = oracle.OLlr(ogroups("mys1080-v4"), 4, L, L, 1, b2, e * nplaq, 0.0225 * nplaq, 1.2 + (-0.7 - e) * 1.125, seed,
                    start=start, K=K, NRM=NRM, cap=cap)
    for _ in range(cap + K * NRM + 2):
        o.step(1)
        if o.status()["phase"] == 2:
            break
    return o.status()


def active_footprint(admit, done, fps, t):
    return sum(f for a, d, f in zip(admit, done, fps) if a <= t < d)


def test_serving_admission_by_footprint(dsa, dgroups, oracle, ogroups):
    """Five requests arriving together (one 12^4, four 8^4) into 8 slots with
    an HBM budget of one 12^4 plus one 8^4 chain: the first two are admitted
    at once, the third waits for a retirement, admission follows arrival
    order, the admitted footprints never exceed the budget, and every
    result equals the oracle's replay of that request alone."""
    g = dgroups("mys1080-v4")
    params = [served_params(dsa, *r) for r in SERVED]
    fps = [dsa.dsa.llr_footprint(p) for p in params]
    budget = fps[0] + fps[1]
    assert fps[0] + 2 * fps[1] > budget
    say(f"case serving: footprints {fps}, budget {budget}")
    with dsa.dsa.LlrEngine(4, 8, hbm_budget=budget) as eng:
        admit, done, st = eng.serve([g] * len(params), params, [0.0] * len(params), chunk=2)
    say("  admit", np.round(admit, 4).tolist(), "done", np.round(done, 4).tolist())
    assert all(s["phase"] == 2 for s in st)
    assert np.all(np.diff(admit) >= 0), "admission out of arrival order"
    assert admit[2] >= min(done[0], done[1]), "third request admitted before a retirement freed budget"
    for t in admit:
        assert active_footprint(admit, done, fps, t) <= budget
    for r, req in enumerate(SERVED):
        want = oracle_served(oracle, ogroups, *req)
        for k in ("a", "S1", "S2", "last_da", "phase", "iterations", "seed_sweeps", "reached", "sweeps", "hit",
                  "acc"):
            assert st[r][k] == want[k], (r, k, st[r][k], want[k])


def test_serving_rejects_request_beyond_budget(dsa, dgroups, oracle, ogroups):
    """A 16^4 request whose footprint exceeds the whole budget is rejected on
    arrival (phase "rejected"); the requests around it are served as alone,
    and the engine serves again afterwards."""
    g = dgroups("mys1080-v4")
    reqs = [SERVED[1], (16, 0.0, -0.4, "hot", 46), SERVED[3]]
    params = [served_params(dsa, *r) for r in reqs]
    fps = [dsa.dsa.llr_footprint(p) for p in params]
    budget = 2 * fps[0]
    assert fps[1] > budget
    say("case serving rejection: footprints", fps, "budget", budget)
    with dsa.dsa.LlrEngine(4, 4, hbm_budget=budget) as eng:
        admit, done, st = eng.serve([g] * 3, params, [0.0, 0.0, 0.0], chunk=2)
        assert st[1]["phase"] == 4 and st[1]["sweeps"] == 0
        assert done[1] == admit[1]
        for r in (0, 2):
            want = oracle_served(oracle, ogroups, *reqs[r])
            assert st[r]["phase"] == 2
            assert (st[r]["a"], st[r]["S1"], st[r]["sweeps"]) == (want["a"], want["S1"], want["sweeps"])
        _, _, st2 = eng.serve([g], params[:1], [0.0], chunk=2)
        assert st2[0]["a"] == st[0]["a"]


@pytest.mark.parametrize("beta,seed,start", [((0.0, 3.0, -0.25), 11, "hot"), ((0.0, 3.5, 0.0), 11, "cold")],
                         ids=["hot-3.0--0.25", "cold-3.5-0"])
def test_benchmark_workload_class_equals_table(dsa, dgroups, beta, seed, start):
    """32^3 x 64 S(1080), three sweeps: the default (class) kernel's links and
    hit/acc counters equal the table kernel's bit for bit.  The table kernel
    computes every trial action from element ReTr values in the reference's
    order (dym-mod-metro.cpp:164-192) and is pinned to the oracle on smaller
    lattices; the class kernel re-decides from element values every hit its
    bound cannot certify (counted and printed).  The cold chain is frozen, so
    its second sweep call runs the class kernel's plaquette-first test."""
    g = dgroups("mys1080-v4")
    out = {}
    for kernel in ("table", "class-narrow", "auto"):
        with dsa.Chain(g, 4, 64, 32, *beta, seed, start=start) as c:
            c.set_kernel(kernel)
            say(f"case 32^3x64 {start} beta={beta}: {kernel} kernel ({c.kernel_form()}), 3 sweeps")
            c.update(1)
            c.update(2)  # below 1e-3 acceptance the second call runs the plaquette-first test
            out[kernel] = (c.a, c.counters(), c.exact_decisions(), c.kernel_form())
    a_t, cnt_t, _, _ = out["table"]
    assert cnt_t[0] == 3 * 20 * 4 * 64 * 32 ** 3
    for kernel in ("class-narrow", "auto"):
        a_c, cnt_c, ex, form = out[kernel]
        say(f"  counters {cnt_t}; acceptance {cnt_t[1] / cnt_t[0]:.5f}; {kernel} runs {form}: "
            f"{ex} hits re-decided from element values ({ex / cnt_t[0]:.2e} of the hits)")
        # the default takes the class update with GF(4) products for S(1080),
        # and for a frozen chain the class update with table gathers
        assert form == (("field" if start == "hot" else "class-narrow") if kernel == "auto" else kernel)
        assert cnt_c == cnt_t
        assert np.array_equal(a_c, a_t)


def test_field_two_links_per_lane_equals_table(dsa, dgroups):
    """40^3 x 48 S(1080), hot, two sweeps: each pass has 384,000 links, more
    than one 512-lane block per CU can take one per lane and not a multiple
    of 1024, so the field kernel runs 375 blocks with one or two links per
    lane (the grid-stride tail).  Its links and counters equal the table
    kernel's bit for bit."""
    g = dgroups("mys1080-v4")
    out = {}
    for kernel in ("table", "field"):
        with dsa.Chain(g, 4, 48, 40, 0.0, 3.0, -0.25, 5) as c:
            c.set_kernel(kernel)
            say(f"case 40^3x48 hot: {kernel} kernel, 2 sweeps")
            c.update(2)
            out[kernel] = (c.a, c.counters())
    assert out["field"][1] == out["table"][1]
    assert np.array_equal(out["field"][0], out["table"][0])
