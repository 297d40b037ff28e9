"""Configuration and manifest files of the reference's workflows.

* raw configurations: the int32 link array a[i*D + mu] in host byte order,
  V*D entries, no header (llr2d.cpp:262-271 save_cfg / load_cfg);
* the llr2d seed manifest: one "CELL: idx i j E1 E2 hw1 hw2 path" line per
  seeded cell (llr2d.cpp:451, read back with the sscanf of llr2d.cpp:500);
* ANE / ANE2 result lines (llr.cpp:250, llr2d.cpp:505);
* SU(3) configurations in the Kentucky byte order written by
  dym-mod-metro-savecfg.cpp:55-86 (the writer itself runs on the device:
  Chain.save_kentucky); read_kentucky reads them back for analysis.
"""
import os
import re

import numpy as np


def write_cfg(path, a):
    """Raw int32 configuration (llr2d.cpp:262-265)."""
    np.ascontiguousarray(a, dtype=np.int32).tofile(path)


def read_cfg(path, V, D):
    """Raw int32 configuration of V*D links; raises unless the file holds
    exactly that many (llr2d.cpp:266-271 returns false then)."""
    a = np.fromfile(path, dtype=np.int32)
    if a.size != V * D:
        raise ValueError(f"{path}: {a.size} int32 entries, expected {V * D}")
    return a


def cell_line(idx, i, j, E1, E2, hw1, hw2, path):
    """One manifest line as llr2d's seed pass prints it (llr2d.cpp:451)."""
    return f"CELL: {idx:d} {i:d} {j:d} {E1:.2f} {E2:.2f} {hw1:.2f} {hw2:.2f} {path}"


_CELL = re.compile(r"^CELL:\s*(-?\d+)\s+(-?\d+)\s+(-?\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)")


def parse_manifest(lines):
    """Cells of a manifest (llr2d.cpp:496-501): lines not starting with
    "CELL:" and lines that do not hold all eight fields are skipped."""
    out = []
    for line in lines:
        if not line.startswith("CELL:"):
            continue
        m = _CELL.match(line)
        if not m:
            continue
        try:
            idx, i, j = (int(m.group(k)) for k in (1, 2, 3))
            E1, E2, hw1, hw2 = (float(m.group(k)) for k in (4, 5, 6, 7))
        except ValueError:
            continue
        out.append({"idx": idx, "i": i, "j": j, "E1": E1, "E2": E2, "hw1": hw1, "hw2": hw2,
                    "path": m.group(8)[:1023]})
    return out


def read_manifest(path):
    with open(path) as f:
        return parse_manifest(f.read().splitlines())


def ane_line(Ewin, mean, sd, n):
    """llr.cpp's result line for one window: mean and standard deviation of
    a_n over its n repeats (llr.cpp:335-339); n = 0 prints the unreached form."""
    if n == 0:
        return f"ANE: {Ewin:8.2f}   --(unreached)"
    return f"ANE: {Ewin:8.2f}  {mean:9.5f}  {sd:9.5f}  {n:d}"


def ane2_line(E1, E2, a1, a2, e1=0.0, e2=0.0, n=1):
    """llr2d.cpp:505's result line for one cell."""
    return f"ANE2: {E1:9.2f} {E2:9.2f} {a1:10.6f} {a2:10.6f} {e1:9.5f} {e2:9.5f} {n:d}"


def kentucky_name(prefix, beta1, beta2, Nt, Nx, num):
    """File name of savekyconfig (dym-mod-metro-savecfg.cpp:78-80)."""
    return f"{prefix}lat-beta1{beta1:.3f}-beta2{beta2:.3f}-nt{Nt:02d}-nx{Nx:02d}-num{num:04d}"


def read_kentucky(path, Nt, Nx):
    """SU(3) links of a Kentucky-order file as U[d, x0, x1, x2, x3] 3x3
    complex matrices (D = 4)."""
    with open(path, "rb") as f:
        return read_kentucky_bytes(f.read(), Nt, Nx)


def read_kentucky_bytes(data, Nt, Nx):
    """read_kentucky on the file's bytes."""
    V = Nt * Nx ** 3
    raw = np.frombuffer(data, dtype=">f8")
    if raw.size != V * 4 * 18:
        raise ValueError(f"{raw.size} doubles, expected {V * 4 * 18}")
    # index x1 + Nx*(x2 + Nx*(x3 + Nx*(x0 + Nt*(part + 2*(r + 3*(c + 3*dd)))))), dd = (d+3)%4
    b = raw.reshape(4, 3, 3, 2, Nt, Nx, Nx, Nx)  # dd, c, r, part, x0, x3, x2, x1
    m = b[..., 0, :, :, :, :] + 1j * b[..., 1, :, :, :, :]  # dd, c, r, x0, x3, x2, x1
    m = np.transpose(m, (0, 3, 6, 5, 4, 2, 1))  # dd, x0, x1, x2, x3, r, c
    return m[[(d + 3) % 4 for d in range(4)]]
