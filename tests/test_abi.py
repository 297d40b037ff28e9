"""The drop-in boundary without a GPU: libdsa.so loads and exports every
symbol of include/dsa.h, the host-side table loader agrees with the oracle,
the visiting order is race-free and equals the oracle's, errors are returned
(never aborts), and the dym-mod-metro CLI keeps the reference's usage contract."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, group_file

HEADER = os.path.join(REPO, "include", "dsa.h")
CLI = os.path.join(REPO, "dym_symanzik_action_amd", "build", "dym-mod-metro")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dsa_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(dsa):
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", dsa.dsa.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (dsa_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert dsa.lib().dsa_version().decode().startswith("dsa-mi355x")


@pytest.mark.parametrize("name", ["mys1080-v4", "S216", "S108", "myBI", "Q8"])
def test_group_loader_matches_oracle(dgroups, ogroups, name):
    g, o = dgroups(name), ogroups(name)
    assert (g.P, g.id, g.npool) == (o.P, o.id, o.npool)
    for a, b in zip(g.tables(), o.tables()):
        assert np.array_equal(a, b)
    if name == "mys1080-v4":
        assert g.min_retr == -3.0 and f"{g.nn_retr:e}" == "-1.618034e+00"


def test_group_loader_errors(dsa, tmp_path):
    with pytest.raises(dsa.DSAError) as e:
        dsa.load_group(str(tmp_path / "nope"))
    assert e.value.code == -2
    bad = tmp_path / "trunc"
    bad.write_text("3\n0 0 0\n0 0 0\n0 1 2 1 2 0 2 0\n")
    with pytest.raises(dsa.DSAError) as e:
        dsa.load_group(str(bad))
    assert e.value.code == -3 and "truncated" in str(e.value)
    noid = tmp_path / "noid"
    noid.write_text("2\n0 0\n0 0\n1 1 1 1\n")
    with pytest.raises(dsa.DSAError) as e:
        dsa.load_group(str(noid))
    assert "identity" in str(e.value)
    rng = tmp_path / "range"
    rng.write_text("2\n0 0\n0 0\n0 1 1 7\n")
    with pytest.raises(dsa.DSAError) as e:
        dsa.load_group(str(rng))
    assert "out of range" in str(e.value)


@pytest.mark.parametrize("shape", [(4, 2, 2), (4, 4, 4), (4, 32, 16), (4, 8, 8), (4, 12, 12), (4, 16, 8),
                                   (4, 8, 16), (4, 3, 3), (4, 6, 4), (3, 6, 4), (2, 8, 6), (4, 5, 4)])
def test_colour_schedule_equals_oracle_and_is_race_free(dsa, oracle, ogroups, shape):
    D, Nt, Nx = shape
    passes = dsa.schedule(D, Nt, Nx, "colour")
    oc = oracle.OChain(ogroups("S108"), D, Nt, Nx, 0, 1, 0, 1, rng="philox", order="colour")
    col, ncol, conflicts = oc.colouring()
    assert conflicts == 0
    V = oc.V
    # every link exactly once per sweep, passes in (mu, colour) order
    seen = np.zeros((D, V), np.int32)
    q = 0
    for mu in range(D):
        for c in range(int(ncol[mu])):
            want = np.nonzero(col[mu] == c)[0]
            if want.size == 0:
                continue
            pmu, sites = passes[q]
            q += 1
            assert pmu == mu and np.array_equal(sites, want)
            seen[mu, sites] += 1
    assert q == len(passes) and np.all(seen == 1)


@pytest.mark.parametrize("shape", [(4, 2, 2), (4, 4, 4), (4, 6, 4), (4, 3, 3), (3, 5, 4)])
def test_reference_schedule_reproduces_sequential_order(dsa, oracle, ogroups, shape):
    """Levels: within a (mu, parity) block no two links of one level share a
    loop, and a link's level exceeds that of every earlier conflicting link."""
    D, Nt, Nx = shape
    passes = dsa.schedule(D, Nt, Nx, "reference")
    oc = oracle.OChain(ogroups("S108"), D, Nt, Nx, 0, 1, 0, 1, rng="minstd", order="reference")
    V = oc.V
    pos = np.stack(np.unravel_index(np.arange(V), [Nx] * (D - 1) + [Nt])[::-1])
    par = pos.sum(0) & 1
    order = {mu: [] for mu in range(D)}
    for mu, sites in passes:
        order[mu].append(sites)
    for mu in range(D):
        level = np.full(V, -1)
        for lv, sites in enumerate(order[mu]):
            level[sites] = lv
        assert np.all(level >= 0)
        for i in range(V):
            rs = oc.readset(i, mu)
            same = np.unique(rs[(rs % D == mu) & (rs // D != i)] // D)
            same = same[par[same] == par[i]]
            assert np.all(level[same] != level[i])
            # parity 0 passes come first, then parity 1 (dym-mod-metro.cpp:153)
            earlier = same[same < i]
            assert np.all(level[earlier] < level[i])
        p0 = [lv for lv, s in enumerate(order[mu]) if par[s[0]] == 0]
        p1 = [lv for lv, s in enumerate(order[mu]) if par[s[0]] == 1]
        assert not p1 or not p0 or max(p0) < min(p1)


def test_chain_without_device_fails_cleanly(dsa, dgroups):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(dsa.DSAError) as e:
        dsa.Chain(dgroups("S108"), 4, 4, 4, 0, 1, 0, 1)
    assert e.value.code == -4


def test_chain_rejects_bad_shapes(dsa, dgroups):
    for args in [(1, 4, 4), (6, 4, 4), (4, 0, 4), (4, 4096, 4), (2, 8, 5000)]:
        with pytest.raises(dsa.DSAError) as e:
            dsa.Chain(dgroups("S108"), *args, 0, 1, 0, 1)
        assert e.value.code in (-1, -5)


def test_cli_usage_and_errors():
    r = subprocess.run([CLI, "a", "4"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("usage:") and r.stdout == ""
    r = subprocess.run([CLI, "/nonexistent", "4", "2", "2", "0", "1", "0", "1", "1", "1"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "cannot open" in r.stderr
    assert r.stdout.startswith("PARAMS(grp,D,Nt,Nx,beta0,beta1,beta2,seed): /nonexistent 4 2 2 ")
    r = subprocess.run([CLI, group_file("Q8"), "4", "2", "2", "0", "1", "0", "1", "--rng", "bogus"],
                       capture_output=True, text=True)
    assert r.returncode == 1


def test_llr_cli_usage_and_errors(tmp_path):
    """llr / llr2d (one binary, mode by name and first argument): short
    argument lists print the reference's usage on stderr and exit 1
    (llr.cpp:278-283, llr2d.cpp:297-304, 420-423, 478-481); a solve pass
    without its manifest exits 1 (llr2d.cpp:490)."""
    llr = os.path.join(REPO, "dym_symanzik_action_amd", "build", "llr")
    llr2d = os.path.join(REPO, "dym_symanzik_action_amd", "build", "llr2d")
    r = subprocess.run([llr, "g", "4"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith(f"usage: {llr} group D Nt Nx axis bg") and r.stdout == ""
    r = subprocess.run([llr2d, "g"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith(f"usage: {llr2d} group D Nt Nx E1top E1bot")
    r = subprocess.run([llr2d, "seed", "g"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("usage: seed seed group D Nt Nx")
    r = subprocess.run([llr2d, "solve", "g"], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("usage: solve solve group D Nt Nx manifest lo hi a0 K NRM seed")
    r = subprocess.run([llr2d, "solve", group_file("Q8"), "4", "3", "3", str(tmp_path / "none.txt"), "0", "1", "1.0",
                        "2", "2", "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "solve: cannot open manifest" in r.stderr
