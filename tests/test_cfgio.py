"""Configuration files and result lines against the reference's own outputs
(tests/golden/llr2d_seed_3x3.npz: an llr2d seed pass's manifest, cell
configurations and a solve pass's ANE2 lines; llr_*.txt: llr's ANE lines;
kentucky_c1.npz: savecfg's SU(3) files).  No device is needed."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.filterwarnings("ignore")


@pytest.fixture(scope="module")
def l2d():
    return np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))


def test_manifest_lines_round_trip(l2d):
    from dym_symanzik_action_amd import cfgio

    lines = [str(l) for l in l2d["manifest"]]
    cells = cfgio.parse_manifest(["# comment", "LLR2D(seed): header"] + lines + ["CELL: 1 2 broken"])
    assert len(cells) == len(lines)
    for c, l in zip(cells, lines):
        assert cfgio.cell_line(c["idx"], c["i"], c["j"], c["E1"], c["E2"], c["hw1"], c["hw2"], c["path"]) == l
    assert [c["idx"] for c in cells] == list(range(len(lines)))


def test_result_lines_match_reference_formats(l2d):
    from dym_symanzik_action_amd import cfgio

    for l in l2d["ane2"]:
        f = str(l).split()
        assert cfgio.ane2_line(float(f[1]), float(f[2]), float(f[3]), float(f[4]), float(f[5]), float(f[6]),
                               int(f[7])) == str(l)
    n = 0
    for name in ("llr_axis1_3x3.txt", "llr_axis2_4x3.txt", "llr_s216_axis1.txt"):
        for l in open(os.path.join(GOLDEN, name)).read().splitlines():
            if l.startswith("ANE:") and "unreached" not in l:
                f = l.split()
                assert cfgio.ane_line(float(f[1]), float(f[2]), float(f[3]), int(f[4])) == l
                n += 1
    assert n > 0


def test_raw_cfg_files_through_the_c_abi(dsa, tmp_path, l2d):
    """dsa_cfg_save writes exactly llr2d's bytes; dsa_cfg_load reads them back
    and refuses files of the wrong length."""
    import ctypes as C

    E_IO, E_RANGE = -2, -5  # include/dsa.h
    L = dsa.dsa.lib()
    from dym_symanzik_action_amd import cfgio

    for c in cfgio.parse_manifest([str(l) for l in l2d["manifest"]]):
        key = os.path.basename(c["path"])[:-4]
        a = np.ascontiguousarray(l2d[key])
        p = tmp_path / (key + ".cfg")
        assert L.dsa_cfg_save(os.fsencode(p), a.ctypes.data_as(C.c_void_p), a.size) == 0
        assert p.read_bytes() == a.tobytes()
        b = np.zeros_like(a)
        assert L.dsa_cfg_load(os.fsencode(p), b.ctypes.data_as(C.c_void_p), b.size) == 0
        assert np.array_equal(a, b)
        short = np.zeros(a.size + 1, np.int32)
        assert L.dsa_cfg_load(os.fsencode(p), short.ctypes.data_as(C.c_void_p), short.size) == E_RANGE
        assert L.dsa_cfg_load(os.fsencode(p), b.ctypes.data_as(C.c_void_p), b.size - 1) == E_RANGE
    assert L.dsa_cfg_load(os.fsencode(tmp_path / "missing.cfg"), b.ctypes.data_as(C.c_void_p), 1) == E_IO


def test_group_matrices_form_the_defining_representation(dsa, dgroups):
    """The loader keeps the 18 reals per element appended to mys1080-v4: every
    matrix is in SU(3), the map is a homomorphism of the table, and
    -Re tr reproduces the file's ReTr column (the table's sign convention)."""
    g = dgroups("mys1080-v4")
    M = g.matrices()
    mult, inv, retr, imtr, pool = g.tables()
    assert M.shape == (g.P, 3, 3)
    assert np.abs(M @ np.conj(np.transpose(M, (0, 2, 1))) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14
    rng = np.random.default_rng(3)
    a, b = rng.integers(0, g.P, 2000), rng.integers(0, g.P, 2000)
    assert np.abs(M[a] @ M[b] - M[mult[a, b]]).max() < 1e-13
    assert np.abs(-np.trace(M, axis1=1, axis2=2).real - retr).max() < 1e-14
    assert np.abs(M[g.id] - np.eye(3)).max() < 1e-15
    with pytest.raises(dsa.dsa.DSAError):
        dgroups("S216").matrices()


def test_kentucky_files_decode_to_group_elements(dsa, dgroups):
    """savecfg's files (2^4, config 1 after 3 and 6 sweeps) hold, at every
    link of the decoded layout, exactly one element's matrix."""
    from dym_symanzik_action_amd import cfgio

    k = np.load(os.path.join(GOLDEN, "kentucky_c1.npz"))
    M = dgroups("mys1080-v4").matrices().reshape(-1, 9)
    for n in range(int(k["N"])):
        U = cfgio.read_kentucky_bytes(k[f"file{n}"].tobytes(), 2, 2)
        flat = U.reshape(-1, 9)
        d = np.abs(flat[:, None, :] - M[None, :, :]).max(axis=2)
        assert np.all(d.min(axis=1) < 1e-15)
        assert np.all((d < 1e-12).sum(axis=1) == 1)
