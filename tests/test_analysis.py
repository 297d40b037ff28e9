"""The analysis module against what fixes its answers: the Gamma method on
series of known integrated autocorrelation time and on the coverage of its
error bars; the parsers on the reference's own output files (the first GMES
line and final ACC of config 1; the scripts' own field reading); the LLR
reconstructions on densities of states with closed forms (constant and
linear slopes, a Gaussian, a symmetric double well; a quadratic ln rho in 2-D);
llr2d's cell indices recovered from its header against the reference's
manifest."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from dym_symanzik_action_amd import analysis as A
from dym_symanzik_action_amd import cfgio


def ar1(rng, rho, n, m=1):
    """m independent AR(1) series x_t = rho x_{t-1} + e_t, started stationary."""
    e = rng.standard_normal((m, n))
    x = np.empty((m, n))
    x[:, 0] = e[:, 0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[:, t] = rho * x[:, t - 1] + e[:, t]
    return x


def test_gamma_method_uncorrelated():
    x = np.random.default_rng(1).standard_normal(200_000)
    r = A.gamma_method(x)
    assert abs(r["tau_int"] - 0.5) < 0.03
    assert abs(r["err"] / (x.std() / np.sqrt(x.size)) - 1) < 0.05
    assert r["N"] == x.size


def test_gamma_method_ar1_tau_int():
    """tau_int of AR(1) is (1 + rho) / (2 (1 - rho)) = 4.5 at rho = 0.8; the
    error of the mean is sigma_x sqrt(2 tau_int / N)."""
    rho = 0.8
    x = ar1(np.random.default_rng(2), rho, 400_000)[0]
    r = A.gamma_method(x)
    exact = (1 + rho) / (2 * (1 - rho))
    assert abs(r["tau_int"] - exact) < 3 * r["dtau_int"] and r["dtau_int"] < 0.3
    sig = 1 / np.sqrt(1 - rho * rho)
    assert abs(r["err"] / (sig * np.sqrt(2 * exact / x.size)) - 1) < 0.1


def test_gamma_method_error_covers_the_mean():
    """Over 300 independent correlated series (rho = 0.9, tau_int 9.5) the
    true mean 0 lies within one error bar about 68 % of the time; the naive
    error (tau_int = 1/2) would cover about 25 %."""
    xs = ar1(np.random.default_rng(3), 0.9, 4000, 300)
    inside = [abs(r["mean"]) < r["err"] for r in map(A.gamma_method, xs)]
    assert 0.58 < np.mean(inside) < 0.78


def test_gamma_method_constant_and_empty():
    r = A.gamma_method(np.full(10, 2.5))
    assert (r["mean"], r["err"], r["tau_int"]) == (2.5, 0.0, 0.5)
    with pytest.raises(ValueError):
        A.gamma_method([])


def test_read_metro_on_reference_output():
    """tests/golden/cli_c1.txt is the reference dym-mod-metro on config 1:
    its first GMES line and final ACC are known answers, and the plaquette
    is what the reference scripts read (field [4] of the line split on
    single spaces, scripts/getavgsimpplaq.py:27-32)."""
    path = os.path.join(GOLDEN, "cli_c1.txt")
    m = A.read_metro(path)
    assert m["params"]["Nt"] == 2 and m["params"]["beta"] == (0.0, 1.5, 0.1) and m["run"] == (10, 100, 100, 20)
    assert m["gmes"].shape == (100, 8) and m["rmes"].shape == (100, 2) and m["wilson"].shape == (100, 2, 2)
    assert np.array_equal(m["gmes"][0], [9.99e2, 0.0, -6.162976e-33, -2.769672, -2.654508, -2.309017, -2.377596,
                                         -1.884393])
    assert m["acc"] == 0.002063
    lines = [l for l in open(path) if "GMES: " in l]
    assert np.array_equal(A.simple_plaquette(m), [float(l.split(" ")[4]) / 3 + 1 for l in lines])
    assert np.array_equal(A.polyakov_modulus(m), np.hypot(m["gmes"][:, 1], m["gmes"][:, 2]))
    assert A.read_metro(open(path).read())["gmes"].shape == (100, 8)


def test_read_metro_timing_lines():
    text = ("TM_BEGIN: update                                   ...\n"
            "TM_END:   update                                   ... done in 1.250e-02 sec.\n")
    assert A.read_metro(text)["timing"] == [("update", 0.0125)]


def test_read_llr_on_reference_output():
    d = A.read_llr(os.path.join(GOLDEN, "llr_axis1_3x3.txt"))
    assert list(d["E"]) == [-800.0, -650.0, -500.0]
    assert list(d["a"]) == [-0.85094, -0.85133, -0.85128] and list(d["nrep"]) == [2, 2, 2]
    assert (d["nplaq"], d["nrect"], d["axis"], d["norm"]) == (486, 972, 1, 486)
    text = "NPLAQ: 96\nNRECT: 192\nAXIS: 2\n" + cfgio.ane_line(-10, 0.5, 0.1, 3) + "\n" + cfgio.ane_line(-20, 0, 0, 0)
    d = A.read_llr(text)
    assert d["norm"] == 192 and d["unreached"] == [-20.0] and list(d["E"]) == [-10.0]


def test_llr_log_density_exact_for_linear_slopes():
    E = np.linspace(-600.0, -300.0, 13)
    Ef, af, ln = A.llr_log_density(E, np.full(E.size, 1.7))
    assert np.allclose(ln, 1.7 * (Ef - E[0]), rtol=0, atol=1e-9)
    al, be = 2e-3, 0.4
    Ef, af, ln = A.llr_log_density(E, al * E + be)
    exact = 0.5 * al * (Ef ** 2 - E[0] ** 2) + be * (Ef - E[0])
    assert np.allclose(ln, exact, rtol=0, atol=1e-9)
    assert set(E) <= set(Ef)
    with pytest.raises(ValueError):
        A.llr_log_density(E[::-1], E)


def test_llr_canonical_gaussian():
    """ln rho = -(E - mu)^2 / (2 s^2) + beta E: P_beta is the Gaussian (mu, s^2)."""
    mu, s, beta = -400.0, 12.0, 1.3
    E = np.linspace(mu - 12 * s, mu + 12 * s, 49)
    Ef, af, ln = A.llr_log_density(E, -(E - mu) / s ** 2 + beta, npts=20001)
    p, mean, var = A.llr_canonical(Ef, ln, beta)
    assert abs(mean - mu) < 1e-6 * s and abs(var / s ** 2 - 1) < 1e-4
    _, mean2, _ = A.llr_canonical(Ef, ln, beta + 0.05)  # larger coupling: lower <E>
    assert abs(mean2 - (mu - 0.05 * s * s)) < 1e-3 * s


def test_llr_transition_symmetric_double_well():
    """ln rho = b_c E - k (E^2 - 1)^2: at beta = b_c the two maxima of
    ln rho - beta E sit at E = -1, +1 with equal height; the slopes are
    sampled at window centres 0.05 apart, as a ladder would give them."""
    bc, k = 2.2, 30.0
    E = np.arange(-1.8, 1.8001, 0.05)
    a = bc - 4 * k * E * (E * E - 1)
    Ef, af, ln = A.llr_log_density(E, a, npts=36001)
    t = A.llr_transition(Ef, ln, np.linspace(a.min(), a.max(), 600))
    assert t is not None
    assert abs(t["beta_c"] - bc) < 2e-3
    assert abs(t["E_low"] + 1) < 2e-3 and abs(t["E_high"] - 1) < 2e-3 and abs(t["latent"] - 2) < 4e-3
    # one phase only: no transition
    assert A.llr_transition(Ef, bc * Ef - (Ef ** 2), np.linspace(0, 4, 50)) is None


def test_read_llr2d_indices_match_reference_manifest():
    """The reference's own llr2d output (the seed header and the solve pass's
    ANE2 lines of tests/golden/llr2d_seed_3x3.npz): the cells' (i, j)
    recovered from the header geometry are the manifest's."""
    f = np.load(os.path.join(GOLDEN, "llr2d_seed_3x3.npz"))
    head = str(f["head"][0])
    text = "\n".join([head] + [str(l) for l in f["ane2"]]) + "\n"
    cells, geo = A.read_llr2d(text)
    assert (geo["E1top"], geo["E1bot"], geo["step1"], geo["nperp"]) == (-400.0, -600.0, 100.0, 1.0)
    man = cfgio.parse_manifest([str(l) for l in f["manifest"]])
    assert A.llr2d_grid_index(cells, geo) == [(c["i"], c["j"]) for c in man]
    assert cells.shape == (len(man), 7)


def test_llr2d_log_density_exact_for_quadratic():
    """Gradients of a quadratic ln rho at the cell centres of a curved band
    (with one cell missing) integrate back to ln rho exactly, up to the
    gauge at the cell of largest E1."""
    geo = {"E1top": -400.0, "E1bot": -700.0, "step1": 50.0, "c0": 10.0, "c1": 0.8, "c2": 1e-4, "step2": 30.0,
           "nperp": 2}
    A2, B2, C2, D2, E2c = 1e-4, -3e-5, 2e-4, 1.1, 0.6

    def lnr(x, y):
        return A2 * x * x + B2 * x * y + C2 * y * y + D2 * x + E2c * y

    rows = []
    for i in range(7):
        for j in range(5):
            if (i, j) == (3, 1):
                continue
            x = geo["E1top"] - i * geo["step1"]
            y = geo["c0"] + geo["c1"] * x + geo["c2"] * x * x + (j - 2) * geo["step2"]
            rows.append([x, y, 2 * A2 * x + B2 * y + D2, B2 * x + 2 * C2 * y + E2c, 0, 0, 1])
    cells = np.array(rows)
    idx = A.llr2d_grid_index(cells, geo)
    assert len(set(idx)) == len(idx) == 34
    ln = A.llr2d_log_density(cells, idx)
    exact = lnr(cells[:, 0], cells[:, 1])
    ref = max(range(len(cells)), key=lambda k: (cells[k, 0], cells[k, 1]))
    assert np.allclose(ln, exact - exact[ref], rtol=0, atol=1e-8)
    lev, F = A.llr2d_free_energy(cells, ln, 0.3, 0.2)
    assert list(lev) == sorted(set(cells[:, 0]))
    m = cells[:, 0] == lev[0]
    x = ln[m] - 0.2 * cells[m, 1]
    assert np.isclose(F[0], -0.3 * lev[0] + np.log(np.exp(x).sum()))


def test_cli_summaries(capsys):
    assert A._main(["metro", os.path.join(GOLDEN, "cli_c1.txt")]) == 0
    out = capsys.readouterr().out
    assert out.startswith("plaquette S/3+1") and "tau_int" in out and "acceptance       0.002063" in out
    assert A._main(["llr", os.path.join(GOLDEN, "llr_axis1_3x3.txt")]) == 0
    assert "3 windows reached" in capsys.readouterr().out
    assert A._main(["llr", os.path.join(GOLDEN, "llr_stats_4x4.txt")]) == 1  # one window: nothing to integrate
    assert A._main([]) == 1
