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


def z_limit(m, z_family=2.0):
    """Largest |z| allowed for each of m comparisons made together, so that a
    correct sampler fails the family no more often than one two-sided
    z_family test (Bonferroni): the normal quantile at 1 - alpha / (2 m),
    alpha = 2 (1 - Phi(z_family)).  m = 1 gives z_family itself; for
    z_family = 2: 2.28 for 2, 2.77 for 8, 2.90 for 12, 3.02 for 18."""
    from statistics import NormalDist

    nd = NormalDist()
    alpha = 2.0 * (1.0 - nd.cdf(z_family))
    return nd.inv_cdf(1.0 - alpha / (2.0 * m))


def p_limit(m, z_family=2.0):
    """The per-comparison significance level of m comparisons made together
    at the family-wise level of one two-sided z_family test (Bonferroni):
    alpha / m, alpha = 2 (1 - Phi(z_family)) = 0.0455 for z_family = 2."""
    from statistics import NormalDist

    return 2.0 * (1.0 - NormalDist().cdf(z_family)) / m


def welch_p(m1, s1, n1, m2, s2, n2):
    """Two-sided p-value of Welch's t-test that two means, each estimated from
    n independent replicas with sample standard deviation s, are equal (the
    Welch-Satterthwaite degrees of freedom; small replica counts make the
    t distribution's tails, not the normal's, the right reference)."""
    from scipy.stats import t as tdist

    v1, v2 = s1 * s1 / n1, s2 * s2 / n2
    t = abs(m1 - m2) / np.sqrt(v1 + v2)
    dof = (v1 + v2) ** 2 / (v1 * v1 / (n1 - 1) + v2 * v2 / (n2 - 1))
    return float(2.0 * tdist.sf(t, dof))


def welch_p_samples(x, y):
    x, y = np.asarray(x, float), np.asarray(y, float)
    return welch_p(x.mean(), x.std(ddof=1), x.size, y.mean(), y.std(ddof=1), y.size)
