"""dsa_chain_measure_noalloc (the measurement entry point that allocates
nothing: its work space is the library's static per-GPU array) against
dsa_chain_measure, through the C ABI: bit-identical on two calls in a row,
the second with a smaller lattice and grid, in both summation modes; and a
lattice whose work space exceeds the array is refused with DSA_E_RANGE."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_measure_equals_measure(dsa, dgroups):
    g = dgroups("mys1080-v4")
    say("case no-alloc measure: 16^3x32 with an 8x8 grid, then 4^4 with 2x2")
    with dsa.Chain(g, 4, 32, 16, 0.0, 3.0, -0.25, 7) as big, dsa.Chain(g, 4, 4, 4, 0.0, 1.2, -0.1, 8) as small:
        big.update(2)
        small.update(3)
        # the first call on fresh chains, before the allocating entry point
        got_big = big.measure_noalloc(8, 8)
        got_small = small.measure_noalloc(2, 2)
        assert np.array_equal(got_big, big.measure(8, 8))
        assert np.array_equal(got_small, small.measure(2, 2))
        # again after more sweeps, the smaller case last, both sum modes
        big.update(1)
        small.update(1)
        assert np.array_equal(big.measure_noalloc(8, 8), big.measure(8, 8))
        # the reference's one-lane summation order (small lattices only)
        for mode in ("tree", "ordered"):
            b = small.measure_noalloc(2, 2, sum=mode)
            assert np.array_equal(b, small.measure(2, 2, sum=mode))
        # plaquette, rectangle and Polyakov only
        assert np.array_equal(small.measure_noalloc(0, 0), small.measure(0, 0))


def test_noalloc_measure_refuses_oversized_work_space(dsa, dgroups):
    g = dgroups("mys1080-v4")
    say("case no-alloc measure: 32^3x64 with an 8x8 grid exceeds the work space")
    with dsa.Chain(g, 4, 64, 32, 0.0, 3.0, -0.25, 11, start="cold") as c:
        with pytest.raises(dsa.DSAError) as e:
            c.measure_noalloc(8, 8)
        assert e.value.code == -5
        # without a Wilson grid the same lattice fits, and equals measure()
        assert np.array_equal(c.measure_noalloc(0, 0), c.measure(0, 0))
