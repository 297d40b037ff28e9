"""Batched multi-chain sweeps: every chain of a batch gives exactly the links
and counters of the same chain run alone, measurements agree with the
single-chain path, and hipGraph-replayed sweep+measure cycles reproduce the
eager sequence on every replay."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


MIX = [
    ("mys1080-v4", 4, 4, 4, 0.0, 3.0, -0.25, 101, "hot", "philox"),
    ("S216", 4, 8, 4, 0.0, 2.0, -0.1, 102, "hot", "philox"),
    ("mys1080-v4", 4, 6, 6, 0.0, 1.5, 0.0, 103, "cold", "philox"),
    ("mys1080-v4", 4, 2, 2, 0.0, 1.5, 0.1, 104, "hot", "minstd"),
    ("S216", 4, 4, 8, 0.1, 2.5, -0.3, 105, "hot", "minstd"),
    ("mys1080-v4", 4, 8, 4, 0.0, 2.2, -0.2, 106, "hot", "philox"),
    ("mys1080-v4", 4, 3, 3, 0.0, 1.0, 0.3, 107, "hot", "philox"),
]


def build(dsa, dgroups, mix, kernel="auto", margin_scale=1.0):
    b = dsa.Batch()
    b.set_kernel(kernel, margin_scale)
    for name, D, Nt, Nx, b0, b1, b2, seed, start, rng in mix:
        b.add(dgroups(name), D, Nt, Nx, b0, b1, b2, seed, start=start, rng=rng)
    return b.finalize()


@pytest.mark.parametrize("kernel", [("auto", 1.0), ("table", 1.0), ("auto", 1e300), ("class-narrow", 1.0),
                                    ("field", 1.0)],
                         ids=["auto", "table", "class-all-exact", "class-narrow", "field"])
def test_batch_equals_lone_chains(dsa, dgroups, kernel):
    """Each chain equals the chain alone, with the class update where it fits
    (auto: with GF(4) products for S(1080) and S(216)), the table update
    everywhere, the class update deciding every hit from element values, and
    the class update with its products from table gathers or from the GF(4)
    form for every chain."""
    say("case batch vs lone chains", kernel)
    nsw = 3
    with build(dsa, dgroups, MIX, *kernel) as b:
        b.sweep(nsw)
        if kernel[1] > 1:
            assert b.exact_decisions() > 0
        hits, accs = b.counters()
        for c, (name, D, Nt, Nx, b0, b1, b2, seed, start, rng) in enumerate(MIX):
            with dsa.Chain(dgroups(name), D, Nt, Nx, b0, b1, b2, seed, start=start, rng=rng) as ch:
                ch.update(nsw)
                assert np.array_equal(b.links(c), ch.a), f"chain {c} differs"
                assert (int(hits[c]), int(accs[c])) == ch.counters()
                m = ch.measure(0, 0)
        got = b.measure()
        for c, (name, D, Nt, Nx, b0, b1, b2, seed, start, rng) in enumerate(MIX):
            with dsa.Chain(dgroups(name), D, Nt, Nx, b0, b1, b2, seed, start=start, rng=rng) as ch:
                ch.update(nsw)
                m = ch.measure(0, 0)
            assert np.allclose(got[c], m[:4], rtol=0, atol=1e-13 * 3)


def test_block_size_does_not_change_results(dsa, dgroups):
    """Every fixed block size (256, 512, 1024 links) gives the tuned batch's
    links and counters: the tuner restores the state it sweeps."""
    say("case batch block sizes")
    with build(dsa, dgroups, MIX) as ref:
        ref.sweep(3)
        links, ms = ref.task_links()
        assert links in (256, 512, 1024) and (ms > 0).all()
        want = [ref.links(c) for c in range(len(MIX))]
        counters = ref.counters()
    for n in (256, 512, 1024):
        b = dsa.Batch()
        b.set_task_links(n)
        for name, D, Nt, Nx, b0, b1, b2, seed, start, rng in MIX:
            b.add(dgroups(name), D, Nt, Nx, b0, b1, b2, seed, start=start, rng=rng)
        with b.finalize() as b:
            assert b.task_links()[0] == n
            b.sweep(3)
            for c in range(len(MIX)):
                assert np.array_equal(b.links(c), want[c]), (n, c)
            assert all(np.array_equal(x, y) for x, y in zip(b.counters(), counters))
    with pytest.raises(dsa.DSAError):
        dsa.Batch().set_task_links(300)


def test_graph_cycles_replay_like_eager(dsa, dgroups):
    say("case graph cycles")
    mix = MIX[:3] + MIX[5:6]
    with build(dsa, dgroups, mix) as eager, build(dsa, dgroups, mix) as graph:
        want = []
        for _ in range(4):
            eager.sweep(2)
            want.append(eager.measure())
        got = graph.run_cycles(2, 2)  # capture + two replays
        got = np.concatenate([got, graph.run_cycles(2, 2)])  # two more replays of the same graph
        assert np.array_equal(got, np.array(want))
        for c in range(len(mix)):
            assert np.array_equal(graph.links(c), eager.links(c))


def test_batch_rejects_mixed_dimensions(dsa, dgroups):
    b = dsa.Batch()
    b.add(dgroups("S216"), 4, 4, 4, 0, 1, 0, 1)
    with pytest.raises(dsa.DSAError):
        b.add(dgroups("S216"), 3, 4, 4, 0, 1, 0, 1)
    b.close()
