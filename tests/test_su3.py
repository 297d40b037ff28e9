"""Premises of the matrix sweep kernel on the S(1080) group file (CPU).

The kernel replaces table gathers by FP32 products of the SU(3) matrices the
group file appends: it needs ReTr(g) = -Re Tr rho(g) and rho(a) rho(b) =
rho(ab), and its decisions are certified with a rounding bound of
2048 x 2^-24 x W on the action change (W = sum of |weights| of the 24
staples).  Here: both identities on the file, and the FP32 error of staple
sums and trial actions on random configurations, well inside that bound.
"""
import gzip

import numpy as np
import pytest

from conftest import group_file


@pytest.fixture(scope="module")
def s1080():
    tok = gzip.open(group_file("mys1080-v4"), "rt").read().split()
    P = int(tok[0])
    retr = np.array(tok[1:1 + P], float)
    mult = np.array(tok[1 + 2 * P:1 + 2 * P + P * P], np.int64).reshape(P, P)
    m = np.array(tok[1 + 2 * P + P * P:1 + 2 * P + P * P + 18 * P], float).reshape(P, 18)
    rho = (m[:, 0::2] + 1j * m[:, 1::2]).reshape(P, 3, 3)
    return P, retr, mult, rho


def test_retr_is_minus_real_trace(s1080):
    P, retr, mult, rho = s1080
    assert np.abs(retr + np.trace(rho, axis1=1, axis2=2).real).max() < 1e-12


def test_matrices_represent_the_table(s1080):
    P, retr, mult, rho = s1080
    rng = np.random.default_rng(5)
    a, b = rng.integers(0, P, 20000), rng.integers(0, P, 20000)
    assert np.abs(rho[a] @ rho[b] - rho[mult[a, b]]).max() < 1e-12
    # unitary: the inverse is the conjugate transpose
    assert np.abs(rho @ rho.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12


def test_fp32_staple_sums_inside_the_bound(s1080):
    """Staple-like chains of five random elements, weighted as plaquettes
    (beta1 = 3) and rectangles (beta2 = -0.25), summed and multiplied by a link
    in complex64; the trial-action change against an element proposal, versus
    the same in float64.  The kernel's bound is 2048 x 2^-24 x W."""
    P, retr, mult, rho = s1080
    rng = np.random.default_rng(9)
    r32 = rho.astype(np.complex64)
    w = np.array([3.0] * 6 + [-0.25] * 18)
    W = np.abs(w).sum()
    worst = 0.0
    for _ in range(300):
        S32 = np.zeros((3, 3), np.complex64)
        S64 = np.zeros((3, 3), np.complex128)
        for k in range(24):
            e = rng.integers(0, P, 3 if k < 6 else 5)
            T32, T64 = r32[e[0]], rho[e[0]]
            for x in e[1:]:
                T32, T64 = T32 @ r32[x], T64 @ rho[x]
            S32 += np.float32(w[k]) * T32
            S64 += w[k] * T64
        b, p = rng.integers(0, P, 2)
        M32, M64 = S32 @ r32[b], S64 @ rho[b]
        d32 = np.trace(M32).real - np.trace(r32[p] @ M32).real
        d64 = np.trace(M64).real - np.trace(rho[p] @ M64).real
        worst = max(worst, abs(float(d32) - d64))
    assert worst < 2048 * 2.0 ** -24 * W / 4
