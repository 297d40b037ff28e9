"""Analysis of the drop-ins' output: Monte Carlo errors and LLR densities of states.

Reads the text the dym-mod-metro, llr and llr2d drop-ins print (the
reference's formats, so files written by the reference read the same) and
computes what the reference's analysis scripts compute from them, with
autocorrelation-corrected errors added:

* ``read_metro``: GMES / RMES / ACC lines of dym-mod-metro
  (dym-mod-metro.cpp:321-329) as arrays; ``simple_plaquette`` is the
  normalised plaquette S/3 + 1 the scripts plot (scripts/getavgsimpplaq.py:32,
  field [4] of the space-split GMES line).
* ``gamma_method``: the mean with its error and integrated autocorrelation
  time, by U. Wolff's Gamma method ("Monte Carlo errors with less errors",
  Comput. Phys. Commun. 156 (2004) 143) with the automatic window, its bias
  correction, and the Madras-Sokal error of tau_int.
* ``read_llr`` / ``llr_log_density`` / ``llr_canonical`` / ``llr_transition``:
  the ANE lines of llr (llr.cpp:332-341) integrated to ln rho(E) (a_n is
  d ln rho / dE at the window centre), the canonical distribution
  P_beta(E) ~ rho(E) exp(-beta E) and the coupling where its two maxima have
  equal height, with the latent heat (scripts/llr_reconstruct.py does the
  same steps).
* ``read_llr2d`` / ``llr2d_log_density`` / ``llr2d_free_energy``: the ANE2
  cells of llr2d (llr2d.cpp:483-498) integrated by least squares over the
  edges of the cell grid to ln rho(S1, S2), and the effective free energy
  along S1 with the S2 direction summed (scripts/llr2d_reconstruct.py).

``python -m dym_symanzik_action_amd.analysis metro|llr|llr2d FILE`` prints
the summaries.  Pure numpy; needs no GPU.
"""
import re
import sys

import numpy as np

__all__ = ["gamma_method", "read_metro", "simple_plaquette", "polyakov_modulus", "read_llr", "llr_log_density",
           "llr_canonical", "llr_transition", "read_llr2d", "llr2d_grid_index", "llr2d_log_density",
           "llr2d_free_energy"]


# ------------------------------------------------------------ Gamma method

def gamma_method(x, S=1.5):
    """Mean of the series x with Wolff's Gamma-method error.

    Returns a dict: mean, err, tau_int (bias-corrected), dtau_int
    (Madras-Sokal), W (the window) and N.  The window is the first W where
    g(W) = exp(-W / tau(W)) - tau(W) / sqrt(W N) turns negative, with
    tau(W) = S / ln((2 tau_int(W) + 1) / (2 tau_int(W) - 1)), or the first W
    where tau_int(W) falls to 1/2 or below (no positive autocorrelation
    left to sum)."""
    x = np.asarray(x, np.float64).ravel()
    n = x.size
    if n == 0:
        raise ValueError("gamma_method: empty series")
    mean = float(x.mean())
    d = x - mean
    if n < 2 or not np.any(d):
        return {"mean": mean, "err": 0.0, "tau_int": 0.5, "dtau_int": 0.0, "W": 0, "N": n}
    # Gamma(t) = 1/(N-t) sum_i d_i d_{i+t}, by FFT with zero padding
    f = np.fft.rfft(d, 2 * n)
    gam = np.fft.irfft(f * np.conj(f), 2 * n)[:n] / np.arange(n, 0, -1)
    g0 = gam[0]
    W = 0
    csum = g0
    for w in range(1, n):
        csum += 2.0 * gam[w]
        tau = csum / (2.0 * g0)
        if tau <= 0.5:
            W = w
            break
        tw = S / np.log((2.0 * tau + 1.0) / (2.0 * tau - 1.0))
        if np.exp(-w / tw) - tw / np.sqrt(w * n) < 0:
            W = w
            break
    else:
        W = n - 1
    C = g0 + 2.0 * gam[1:W + 1].sum()
    C *= 1.0 + (2.0 * W + 1.0) / n  # bias of the estimated mean in Gamma(t)
    tau = C / (2.0 * g0)
    return {"mean": mean, "err": float(np.sqrt(max(C, 0.0) / n)), "tau_int": float(tau),
            "dtau_int": float(tau * np.sqrt((4.0 * W + 2.0) / n)), "W": int(W), "N": n}


# ----------------------------------------------------------- dym-mod-metro

def _lines(src):
    """A list of lines, a text (holding a newline) or a path, as lines."""
    if isinstance(src, (list, tuple)):
        return [str(l) for l in src]
    if isinstance(src, str) and "\n" in src:
        return src.splitlines()
    with open(src) as f:
        return f.read().splitlines()


_TM = re.compile(r"done in ([-+\d.eE]+) sec")


def read_metro(src):
    """dym-mod-metro output (a path, the text, or its lines) as arrays.

    Returns a dict: params (PARAMS line: group, D, Nt, Nx, beta0..2, seed),
    run (K, N, Ntherm, NHIT), gmes (N x (4 + Nx Nt): placeholder, Re and Im
    of the Polyakov loop, plaquette, then the Wilson grid k-major), wilson
    (N x Nx x Nt), rmes (N x 2: plaquette, rectangle), acc (the last ACC
    value) and the timing lines as (label, seconds)."""
    out = {"params": None, "run": None, "gmes": [], "rmes": [], "acc": None, "timing": []}
    for l in _lines(src):
        if l.startswith("PARAMS("):
            t = l.split(":", 1)[1].split()
            out["params"] = {"group": t[0], "D": int(t[1]), "Nt": int(t[2]), "Nx": int(t[3]),
                             "beta": (float(t[4]), float(t[5]), float(t[6])), "seed": int(t[7])}
        elif l.startswith("RUN("):
            out["run"] = tuple(int(v) for v in l.split(":", 1)[1].split())
        elif l.startswith("GMES:"):
            out["gmes"].append([float(v) for v in l.split()[1:]])
        elif l.startswith("RMES:"):
            out["rmes"].append([float(v) for v in l.split()[1:3]])
        elif l.startswith("ACC:"):
            out["acc"] = float(l.split()[1])
        elif l.startswith("TM_END:"):
            m = _TM.search(l)
            if m:  # timer.cpp:24: "TM_END:   <label> <label>... done in %.3e sec."
                out["timing"].append((l.split()[1], float(m.group(1))))
    g = np.array(out["gmes"], np.float64) if out["gmes"] else np.zeros((0, 4))
    out["gmes"] = g
    out["rmes"] = np.array(out["rmes"], np.float64).reshape(-1, 2)
    if out["params"] and g.size:
        nx, nt = out["params"]["Nx"], out["params"]["Nt"]
        ngrid = g.shape[1] - 4
        if ngrid == nx * nt:
            out["wilson"] = g[:, 4:].reshape(len(g), nx, nt)
    return out


def simple_plaquette(m):
    """S/3 + 1 per measurement, S the GMES plaquette (0 cold ... 1 random)."""
    return m["gmes"][:, 3] / 3.0 + 1.0


def polyakov_modulus(m):
    """|<P>| per measurement from the GMES Polyakov fields."""
    return np.hypot(m["gmes"][:, 1], m["gmes"][:, 2])


# --------------------------------------------------------------------- llr

def read_llr(src):
    """llr output: dict with E, a, sd, nrep (reached windows, ascending E),
    unreached (their centres), nplaq, nrect, axis and norm (NPLAQ for axis 1,
    NRECT for axis 2: <E> per loop = S / norm / 3 + 1)."""
    E, a, sd, n, unreached = [], [], [], [], []
    info = {"nplaq": None, "nrect": None, "axis": 1}
    for l in _lines(src):
        if l.startswith("NPLAQ:"):
            info["nplaq"] = int(l.split()[1])
        elif l.startswith("NRECT:"):
            info["nrect"] = int(l.split()[1])
        elif l.startswith("AXIS:"):
            info["axis"] = int(l.split()[1])
        elif l.startswith("ANE:"):
            t = l.split()
            if "unreached" in l:
                unreached.append(float(t[1]))
                continue
            E.append(float(t[1]))
            a.append(float(t[2]))
            sd.append(float(t[3]))
            n.append(int(t[4]))
    o = np.argsort(E)
    out = {"E": np.array(E)[o], "a": np.array(a)[o], "sd": np.array(sd)[o], "nrep": np.array(n, int)[o],
           "unreached": sorted(unreached)}
    out.update(info)
    out["norm"] = info["nrect"] if info["axis"] == 2 else info["nplaq"]
    return out


def llr_log_density(E, a, npts=4001):
    """ln rho on a fine grid over [E[0], E[-1]] from the window slopes a_n at
    the (ascending) centres E: a is interpolated linearly between centres and
    integrated exactly (trapezoids of the piecewise-linear a), ln rho(E[0]) = 0.
    Returns (Ef, af, lnrho)."""
    E = np.asarray(E, np.float64)
    a = np.asarray(a, np.float64)
    if E.size < 2 or np.any(np.diff(E) <= 0):
        raise ValueError("llr_log_density: need >= 2 strictly ascending window centres")
    Ef = np.union1d(np.linspace(E[0], E[-1], npts), E)
    af = np.interp(Ef, E, a)
    lnrho = np.concatenate([[0.0], np.cumsum(0.5 * (af[1:] + af[:-1]) * np.diff(Ef))])
    return Ef, af, lnrho


def llr_canonical(Ef, lnrho, beta):
    """Normalised P_beta(E) ~ rho(E) exp(-beta E) on the grid Ef (the weight
    of the action with coupling beta), with <E> and its variance."""
    Ef = np.asarray(Ef, np.float64)
    g = np.asarray(lnrho, np.float64) - beta * Ef
    w = np.exp(g - g.max())
    dx = np.zeros_like(Ef)  # trapezoid weights
    dx[1:] += 0.5 * np.diff(Ef)
    dx[:-1] += 0.5 * np.diff(Ef)
    Z = np.sum(w * dx)
    p = w / Z
    mean = np.sum(p * Ef * dx)
    var = np.sum(p * (Ef - mean) ** 2 * dx)
    return p, float(mean), float(var)


def _two_peaks(Ef, g):
    """(height difference low-E peak minus high-E peak, E_low, E_high) of g
    split at its deepest interior well, or None when g has no interior well
    deeper than rounding."""
    inner = g[1:-1]
    wells = np.nonzero((inner < g[:-2]) & (inner < g[2:]))[0] + 1
    if wells.size == 0:
        return None
    j = wells[np.argmin(g[wells])]
    lo, hi = np.argmax(g[:j + 1]), j + np.argmax(g[j:])
    # a barrier within rounding of the peaks is no barrier
    if lo == j or hi == j or min(g[lo], g[hi]) - g[j] <= 1e-9 * max(1.0, np.abs(g).max()):
        return None
    return g[lo] - g[hi], Ef[lo], Ef[hi]


def llr_transition(Ef, lnrho, betas):
    """The coupling where P_beta(E) has two maxima of equal height (equal
    free energies of the two phases), by bisection on the height difference
    over the bracket of `betas` (ascending) where it changes sign.  Returns
    dict(beta_c, E_low, E_high, latent = E_high - E_low), or None if no
    bimodal coupling in `betas` brackets a sign change."""
    Ef = np.asarray(Ef, np.float64)
    lnrho = np.asarray(lnrho, np.float64)
    prev = None
    for b in betas:
        r = _two_peaks(Ef, lnrho - b * Ef)
        if r is None:
            prev = None
            continue
        if prev is not None and np.sign(prev[1][0]) != np.sign(r[0]):
            lo, hi = prev[0], b
            rlo = prev[1]
            for _ in range(60):
                mid = 0.5 * (lo + hi)
                rm = _two_peaks(Ef, lnrho - mid * Ef)
                if rm is None:
                    break
                if np.sign(rm[0]) == np.sign(rlo[0]):
                    lo, rlo = mid, rm
                else:
                    hi = mid
            bc = 0.5 * (lo + hi)
            rc = _two_peaks(Ef, lnrho - bc * Ef) or rlo
            return {"beta_c": bc, "E_low": float(rc[1]), "E_high": float(rc[2]), "latent": float(rc[2] - rc[1])}
        prev = (b, r)
    return None


# ------------------------------------------------------------------- llr2d

_E1 = re.compile(r"E1=\[([-+\d.eE]+),([-+\d.eE]+)\]")


def read_llr2d(src):
    """llr2d output (full or solve pass): cells as an array of rows
    (E1, E2, a1, a2, sd1, sd2, nrep) and the grid geometry of the LLR2D header
    (E1top, E1bot, step1, hw1, c0, c1, c2, step2, hw2, nperp) when present."""
    cells, geo = [], {}
    for l in _lines(src):
        if l.startswith("LLR2D"):
            m = _E1.search(l)
            if m:
                geo["E1top"], geo["E1bot"] = float(m.group(1)), float(m.group(2))
            for tok in l.replace("|", " ").split():
                if "=" in tok and not tok.startswith("E1="):
                    k, v = tok.split("=", 1)
                    try:
                        geo[k] = float(v)
                    except ValueError:
                        pass
        elif l.startswith("ANE2:"):
            t = l.split()
            cells.append([float(v) for v in t[1:7]] + [int(t[7])])
    return np.array(cells, np.float64).reshape(-1, 7), geo


def llr2d_grid_index(cells, geo):
    """(i, j) of each cell: i along E1 from E1top in steps of step1, j across
    the ridge E2 = c0 + c1 E1 + c2 E1^2 in steps of step2, offset by nperp."""
    E1, E2 = cells[:, 0], cells[:, 1]
    i = np.rint((geo["E1top"] - E1) / geo["step1"]).astype(int)
    ridge = geo["c0"] + geo["c1"] * E1 + geo["c2"] * E1 * E1
    j = np.rint((E2 - ridge) / geo["step2"]).astype(int) + int(geo["nperp"])
    return list(zip(i.tolist(), j.tolist()))


def llr2d_log_density(cells, idx):
    """ln rho at the cell centres from the gradients (a1, a2) = grad ln rho:
    every pair of grid neighbours (i, j)-(i+1, j) and (i, j)-(i, j+1) gives
    ln rho_b - ln rho_a = (grad_a + grad_b) / 2 . (x_b - x_a) (exact for a
    quadratic ln rho), solved in the least-squares sense with ln rho = 0 at
    the cell of largest E1 (then largest E2)."""
    cells = np.asarray(cells, np.float64)
    n = len(cells)
    pos = {ij: k for k, ij in enumerate(idx)}
    rows, rhs = [], []
    for k, (i, j) in enumerate(idx):
        for nb in ((i + 1, j), (i, j + 1)):
            k2 = pos.get(nb)
            if k2 is None:
                continue
            r = np.zeros(n)
            r[k2], r[k] = 1.0, -1.0
            rows.append(r)
            rhs.append(0.5 * (cells[k, 2:4] + cells[k2, 2:4]) @ (cells[k2, :2] - cells[k, :2]))
    ref = max(range(n), key=lambda k: (cells[k, 0], cells[k, 1]))
    r = np.zeros(n)
    r[ref] = 1.0
    rows.append(r)
    rhs.append(0.0)
    sol, *_ = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)
    return sol


def llr2d_free_energy(cells, lnrho, b1, b2):
    """F(E1) = -b1 E1 + ln sum_j exp(ln rho(E1, E2_j) - b2 E2_j): the weight
    of each E1 level with the cells across the ridge summed.  Returns
    (E1 levels ascending, F)."""
    cells = np.asarray(cells, np.float64)
    lev = np.unique(np.round(cells[:, 0], 6))
    F = []
    for e in lev:
        m = np.round(cells[:, 0], 6) == e
        x = lnrho[m] - b2 * cells[m, 1]
        F.append(-b1 * e + x.max() + np.log(np.exp(x - x.max()).sum()))
    return lev, np.array(F)


# --------------------------------------------------------------------- CLI

def _main(argv):
    if len(argv) < 2 or argv[0] not in ("metro", "llr", "llr2d"):
        print("usage: python -m dym_symanzik_action_amd.analysis metro|llr|llr2d FILE [beta ...]", file=sys.stderr)
        return 1
    kind, path = argv[0], argv[1]
    if kind == "metro":
        m = read_metro(path)
        for name, x in (("plaquette S/3+1", simple_plaquette(m)), ("|P|", polyakov_modulus(m)),
                        ("rectangle", m["rmes"][:, 1] if len(m["rmes"]) else np.zeros(0))):
            if x.size == 0:
                continue
            r = gamma_method(x)
            print(f"{name:16s} {r['mean']: .8f} +- {r['err']:.8f}  tau_int {r['tau_int']:.2f} +- "
                  f"{r['dtau_int']:.2f}  W {r['W']}  N {r['N']}")
        if m["acc"] is not None:
            print(f"acceptance       {m['acc']:.6f}")
        return 0
    if kind == "llr":
        d = read_llr(path)
        if len(d["E"]) < 2:
            print(f"{len(d['E'])} window(s) reached: need two to integrate a_n", file=sys.stderr)
            return 1
        Ef, af, lnrho = llr_log_density(d["E"], d["a"])
        betas = np.linspace(d["a"].min(), d["a"].max(), 400)
        t = llr_transition(Ef, lnrho, betas)
        print(f"{len(d['E'])} windows reached, {len(d['unreached'])} unreached; axis {d['axis']}, norm {d['norm']}")
        if t is None:
            print("no two-phase coupling in the a_n range (crossover)")
        else:
            norm = 3.0 * d["norm"] if d["norm"] else 1.0
            print(f"beta_c {t['beta_c']:.6f}  E_low {t['E_low']:.2f}  E_high {t['E_high']:.2f}  "
                  f"latent {t['latent']:.2f}  (per loop {t['latent'] / norm:.6f})")
        return 0
    cells, geo = read_llr2d(path)
    idx = llr2d_grid_index(cells, geo)
    lnrho = llr2d_log_density(cells, idx)
    for (i, j), c, v in zip(idx, cells, lnrho):
        print(f"cell {i:3d} {j:3d}  E1 {c[0]:9.2f}  E2 {c[1]:9.2f}  ln rho {v: .6f}")
    return 0


if __name__ == "__main__":
    sys.exit(_main(sys.argv[1:]))
