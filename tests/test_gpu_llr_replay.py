"""The device replay of the reference llr's serial chain (dsa_llr_ref_*)
against the reference's own ladders (tests/golden/llrladder_*.npz, written
by oracle/_ref/ref_llr_harness from llr.cpp's seed_window / rm_solve):
RECT_MULT, every window's drive-in and energies, a_n, the drift and the
links after every window, bit for bit; and `llr --replay` against the
reference binary's stdout."""
import glob
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, group_file

pytestmark = pytest.mark.gpu

# the unreached window of llrladder_axis2_3x2 drives for the full 20000-sweep
# cap, which the one-lane replay walks in minutes: that ladder (and the
# pass that stops there) is checked on the CPU oracle only
LADDERS = [p for p in sorted(glob.glob(os.path.join(GOLDEN, "llrladder_*.npz"))) if "axis2_3x2" not in p]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("path", LADDERS, ids=os.path.basename)
def test_replay_reproduces_reference_ladder(dsa, dgroups, path):
    from ladder_util import check_ladder, ladder_args

    f = np.load(path)
    g, D, Nt, Nx, axis, bg, *_rest = ladder_args(f)
    seed = ladder_args(f)[11]
    say("case llr replay", os.path.basename(path))
    with dsa.dsa.LlrRef(dgroups(g), D, Nt, Nx, axis, bg, seed) as r:
        check_ladder(r, f)
