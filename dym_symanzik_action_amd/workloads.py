"""The multi-chain workloads this package serves, defined once for the
benchmark and the tests that check them.

* ``scan_chains()``: a phase-diagram scan of 256 independent chains -- the
  device-side replacement of the reference's one-process-per-point scans
  (evermore.sh:1-23, jobs2run:1-55, each process running
  dym-mod-metro.cpp:199-337).
* ``llr_requests()``: a Poisson stream of LLR window requests -- what
  scripts/llr2d_parallel.sh:28-37 farms out as one process per window, each
  request solving one window as llr.cpp:223-250 does (drive-in, then
  Robbins-Monro for a_n).
"""
from __future__ import annotations

import numpy as np

# (D, Nt, Nx) of the scan's lattices, assigned round-robin: 8^4, 12^4,
# 16^3 x 8 (Nt = 8, Nx = 16) and 8^3 x 16 (Nt = 16, Nx = 8)
SCAN_SHAPES = [(4, 8, 8), (4, 12, 12), (4, 8, 16), (4, 16, 8)]


def scan_chains():
    """256 chains on a 16 x 16 grid of beta1 in [1, 4] x beta2 in [-0.5, 0.3]
    (beta0 = 0), shapes round-robin over SCAN_SHAPES, Philox seeds 100 + i,
    hot starts; chains 0..191 use S(1080) (mys1080-v4), 192..255 S(216).
    Returns [(group name, (D, Nt, Nx), beta1, beta2, seed), ...]."""
    out = []
    for i in range(256):
        b1 = 1.0 + 3.0 * (i // 16) / 15.0
        b2 = -0.5 + 0.8 * (i % 16) / 15.0
        out.append(("S216" if i >= 192 else "mys1080-v4", SCAN_SHAPES[i % 4], b1, b2, 100 + i))
    return out


def scan_links(chains=None) -> int:
    """Links of every chain of a scan, summed (one sweep's link-updates)."""
    chains = scan_chains() if chains is None else chains
    return sum(D * Nt * Nx ** (D - 1) for _, (D, Nt, Nx), *_ in chains)


# S(1080): ReTr ranges over [-3, 1.5], so the plaquette action S1 = sum over
# plaquettes of ReTr ranges over NPLAQ * [-3, 1.5], NPLAQ = 6 V (D = 4)
S1080_RETR_MIN, S1080_RETR_MAX = -3.0, 1.5


def llr_requests(nreq: int, seed: int = 2024, rate: float = 50.0, lattices=(8, 12, 16),
                 beta2s=(0.0, -0.1, -0.25), K: int = 20, NRM: int = 50, tol: float = 1e-4, cap: int = 200):
    """A Poisson stream of `nreq` LLR window requests at `rate` per second.

    Each request asks for a_n of one plaquette-action window (axis 1) of full
    width delta = 1 % of the S1 range, delta = 0.01 * 4.5 * NPLAQ, i.e.
    half width hw = delta / 2 (llr.cpp:236: [Ewin - hw, Ewin + hw]); window
    centres are uniform over the range, [-3 + delta/2, 1.5 - delta/2] *
    NPLAQ; the lattice is L^4 with L uniform over `lattices`, beta2 uniform
    over `beta2s`.  A window below the midpoint of the frozen (-3) and
    disordered (~0) plaquette values starts cold, as llr.cpp's up-ladder
    does (llr.cpp:256-274: cold start, walk up), the others hot (the
    down-ladder's start).  Drive-in is capped at `cap` sweeps (windows above
    ~+1.0 per plaquette are not reached by the greedy drive-in; they retire
    unreached, as llr.cpp's "--(unreached)" windows do); then Robbins-Monro
    with K sweeps per iteration until |delta a| < tol or NRM iterations
    (llr.cpp:234-250).  The initial tilt a0 is a linear guess of the beta at
    which <S1>/NPLAQ equals the window centre.

    Returns (arrival times in s, [LlrParams], [L], [centre / NPLAQ])."""
    from .dsa import llr_params

    rng = np.random.default_rng(seed)
    arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, nreq - 1))])
    span = S1080_RETR_MAX - S1080_RETR_MIN
    params, shapes, centres = [], [], []
    for r in range(nreq):
        L = int(rng.choice(list(lattices)))
        b2 = float(rng.choice(list(beta2s)))
        half = 0.005 * span
        e = float(rng.uniform(S1080_RETR_MIN + half, S1080_RETR_MAX - half))
        nplaq = 6 * L ** 4
        a0 = 1.2 + (-0.7 - e) * (1.8 / 1.6)
        start = "cold" if e < 0.5 * S1080_RETR_MIN else "hot"
        params.append(llr_params(4, L, L, 1, b2, e * nplaq, half * nplaq, a0, 1000 + r, start=start, K=K,
                                 NRM=NRM, tol=tol, cap=cap))
        shapes.append(L)
        centres.append(e)
    return arrivals, params, shapes, centres
