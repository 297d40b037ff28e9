"""Configuration files written from device-resident chains, against the
reference: savecfg's SU(3) files byte for byte (tests/golden/kentucky_c1.npz,
a replay of its run in the reference's streams and order), and raw int32
configurations through save/load."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_kentucky_files_match_savecfg_byte_for_byte(dsa, dgroups, tmp_path):
    k = np.load(os.path.join(GOLDEN, "kentucky_c1.npz"))
    b0, b1, b2 = (float(x) for x in k["args"][3:6])
    K, N = int(k["K"]), int(k["N"])
    print("case kentucky config 1", K, N, flush=True)
    with dsa.Chain(dgroups("mys1080-v4"), 4, 2, 2, b0, b1, b2, int(k["seed"]), rng="minstd", order="reference") as c:
        for n in range(N):
            c.update(K)
            name = c.save_kentucky(str(tmp_path) + "/", n)
            assert os.path.basename(name) == str(k["names"][n])
            assert open(name, "rb").read() == k[f"file{n}"].tobytes(), f"config {n} differs"


def test_kentucky_layout_on_a_larger_lattice(dsa, dgroups, tmp_path):
    """4^3 x 6, Philox chain after 2 sweeps: the file decodes (numpy, big-endian,
    the savecfg index formula) to the matrices of the chain's own links."""
    from dym_symanzik_action_amd import cfgio

    g = dgroups("mys1080-v4")
    M = g.matrices()
    with dsa.Chain(g, 4, 6, 4, 0.0, 2.0, -0.1, 13) as c:
        c.update(2)
        name = c.save_kentucky(str(tmp_path) + "/", 7)
        a = c.a.reshape(6 * 4 ** 3, 4)
    U = cfgio.read_kentucky(name, 6, 4)
    for x0 in range(6):
        for x1 in range(4):
            for x2 in range(4):
                for x3 in range(4):
                    i = x0 + 6 * (x1 + 4 * (x2 + 4 * x3))
                    for d in range(4):
                        assert np.array_equal(U[d, x0, x1, x2, x3], M[a[i, d]])


def test_raw_cfg_save_load_round_trip(dsa, dgroups, tmp_path):
    g = dgroups("mys1080-v4")
    p = str(tmp_path / "cell.cfg")
    with dsa.Chain(g, 4, 4, 4, 0.0, 2.0, 0.0, 3) as c, dsa.Chain(g, 4, 4, 4, 0.0, 2.0, 0.0, 3, start="cold") as d:
        c.update(3)
        c.save_cfg(p)
        assert np.array_equal(np.fromfile(p, np.int32), c.a)
        d.load_cfg(p)
        assert np.array_equal(d.a, c.a)
        np.zeros(10, np.int32).tofile(p)
        with pytest.raises(dsa.dsa.DSAError):
            d.load_cfg(p)
        np.full(4 ** 4 * 4, g.P, np.int32).tofile(p)  # element out of range
        with pytest.raises(dsa.dsa.DSAError):
            d.load_cfg(p)
