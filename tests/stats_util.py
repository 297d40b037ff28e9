"""Autocorrelation-corrected errors for Monte Carlo time series.

Gamma method of U. Wolff, "Monte Carlo errors with less errors", Comput.
Phys. Commun. 156 (2004) 143: normalised autocorrelation rho(t), integrated
time tau_int(W) = 1/2 + sum_{t<=W} rho(t), window W chosen by Wolff's
automatic criterion with S = 1.5; error^2 = 2 tau_int var / N.
"""
import numpy as np


def gamma_method(x, S=1.5):
    x = np.asarray(x, np.float64)
    n = x.size
    m = x.mean()
    d = x - m
    var = d.var()
    if var == 0 or n < 4:
        return m, 0.0, 0.5
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / np.arange(n, 0, -1)
    rho = acf / acf[0]
    tau = 0.5
    for W in range(1, n // 2):
        tau = 0.5 + rho[1: W + 1].sum()
        if tau <= 0:
            tau = 0.5
            break
        g = np.exp(-W / (S * tau)) - S * tau / np.sqrt(W * n) if tau > 0 else -1
        if g < 0:
            break
    err = np.sqrt(2 * tau * var / n)
    return m, err, tau


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
