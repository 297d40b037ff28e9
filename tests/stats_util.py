"""Autocorrelation-corrected errors for Monte Carlo time series: the Gamma
method of the package's analysis module (U. Wolff, Comput. Phys. Commun. 156
(2004) 143, automatic window with S = 1.5), as (mean, error, tau_int)."""
import numpy as np

from dym_symanzik_action_amd.analysis import gamma_method as _gamma


def gamma_method(x, S=1.5):
    r = _gamma(x, S)
    return r["mean"], r["err"], r["tau_int"]


def parse_gmes(lines, nx, nt):
    """Per-measurement arrays from GMES/RMES lines: plaq, rect, |P|, ReP, wilson[k,l]."""
    g = [l.split() for l in lines if l.startswith("GMES:")]
    r = [l.split() for l in lines if l.startswith("RMES:")]
    arr = np.array([[float(v) for v in row[1:]] for row in g])
    rect = np.array([float(row[2]) for row in r])
    out = {
        "plaq": arr[:, 3],
        "rect": rect,
        "absP": np.hypot(arr[:, 1], arr[:, 2]),
        "ReP": arr[:, 1],
    }
    w = arr[:, 4:].reshape(len(arr), nx, nt)
    for k in range(nx):
        for l in range(nt):
            out[f"W{k + 1}x{l + 1}"] = w[:, k, l]
    return out
