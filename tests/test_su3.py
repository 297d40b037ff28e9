"""Premises of the matrix sweep kernel on the S(1080) group file (CPU).

The kernel sums the SU(3) matrices the group file appends, M = sum_k w_k
rho(t_k) over the 24 staple-times-link elements, and takes each trial action
from one FP32 trace: it needs ReTr(g) = -Re Tr rho(g) and rho(a) rho(b) =
rho(ab), and its decisions are certified with a rounding bound of
1024 x 2^-24 x W on the action change (W = sum of |weights| of the 24
staples; the derivation gives < 420 x 2^-24 x W).  Here: both identities on
the file, and the FP32 error of that arithmetic on random elements, inside
the derived bound.
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


def test_fp32_trial_actions_inside_the_bound(s1080):
    """The kernel's arithmetic in float32, one rounding per operation as its
    FMA chains round: M = sum_k w_k rho(t_k) (plaquettes beta1 = 3, rectangles
    beta2 = -0.25), Re Tr M, Re Tr(rho(p) M) as an 18-term chain, then an
    accepted proposal's M rho(p) and the next trial; against float64."""
    P, retr, mult, rho = s1080
    rng = np.random.default_rng(9)
    f = np.float32
    r32 = rho.astype(np.complex64)
    w = np.array([3.0] * 6 + [-0.25] * 18)
    W = np.abs(w).sum()
    u = 2.0 ** -24

    def trace_chain(R, M):  # Re Tr(R M), FMA order of the kernel
        t = f(0)
        for i in range(3):
            for j in range(3):
                t = f(t + f(R[i, j].real) * f(M[j, i].real))
                t = f(t - f(R[i, j].imag) * f(M[j, i].imag))
        return t

    worst0 = worst1 = 0.0
    for _ in range(400):
        t = rng.integers(0, P, 24)
        M32 = np.zeros((3, 3), np.complex64)
        for k in range(24):
            M32 = (M32 + np.complex64(w[k]) * r32[t[k]]).astype(np.complex64)
        M64 = sum(w[k] * rho[t[k]] for k in range(24))
        p, q = rng.integers(0, P, 2)
        v32 = f(f(M32[0, 0].real + M32[1, 1].real) + M32[2, 2].real)
        d32 = float(v32) - float(trace_chain(r32[p], M32))
        d64 = np.trace(M64).real - np.trace(rho[p] @ M64).real
        worst0 = max(worst0, abs(d32 - d64))
        # after accepting p: M rho(p), then a trial of q
        M32 = (M32 @ r32[p]).astype(np.complex64)
        M64 = M64 @ rho[p]
        v32 = f(f(M32[0, 0].real + M32[1, 1].real) + M32[2, 2].real)
        d32 = float(v32) - float(trace_chain(r32[q], M32))
        d64 = np.trace(M64).real - np.trace(rho[q] @ M64).real
        worst1 = max(worst1, abs(d32 - d64))
    print(f"worst {worst0 / (u * W):.1f} and {worst1 / (u * W):.1f} x 2^-24 W")
    assert worst0 < 420 * u * W
    assert worst1 < (420 + 100) * u * W
    # the bound is not vacuous: rounding is visible at this scale
    assert worst0 > 0.1 * u * W
