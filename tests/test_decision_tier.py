"""The single-precision tier of the certified hit decision (CPU).

The field sweep kernel decides each hit first in floats: p1 = b1 d1, p2 =
b2 d2 from the exact integer differences d1, d2 of the class sums, the band
Mf = margin + 2^-21 (|p1| + |p2|) around dcf = p1 + p2, and the float test of
metro_certified_f (device_common.hpp).  It may only decide a hit when every
action change in the class band [dc - margin, dc + margin] gets the same
decision as the reference's test newact <= oldact or exp(-(newact - oldact))
> u (dym-mod-metro.cpp:176-188).  The checks below restate the float tier in
numpy float32 and compare each decision it makes with the exact decisions at
both ends of the band (the decision is monotone in the change), on random
cases and on cases placed within a few float ulps of every threshold."""
import math

import numpy as np

F = np.float32


def tier(d1, d2, b1q, b2q, margin, u2):
    """The float tier: 1 accept, 0 reject, -1 undecided (as the device)."""
    b1f, b2f = F(b1q), F(b2q)
    mf = F(margin) * F(1.0 + 2.0 ** -22)
    p1, p2 = b1f * F(d1), b2f * F(d2)
    dc = p1 + p2
    M = mf + F(2.0 ** -21) * (abs(p1) + abs(p2))
    u = F(u2)
    if dc + M < F(0):
        return 1
    lo = dc - M
    if lo > F(0):
        if lo > F(37):
            return 0 if u >= F(2e-15) else -1
        if dc >= F(60):
            return -1
        e = np.exp2(-dc * F(1.44269504))
        w = F(1.02) * M + F(2.0 ** -16)
        return 1 if u < e * (F(1) - w) else (0 if u >= e * (F(1) + w) else -1)
    return 1 if u < F(1) - (F(2.02) * M + F(2.0 ** -22)) else -1


def exact(x, u2):
    """The reference's decision for an action change x."""
    return 1 if x <= 0 or math.exp(-x) > u2 else 0


def check(d1, d2, b1q, b2q, margin, u2):
    dec = tier(d1, d2, b1q, b2q, margin, u2)
    if dec < 0:
        return 0
    dc = b1q * d1 + b2q * d2  # float64: within 2^-52 |dc| of the exact value
    slack = 4 * 2.0 ** -52 * (abs(b1q * d1) + abs(b2q * d2))
    for x in (dc - margin - slack, dc + margin + slack):
        assert exact(x, u2) == dec, (d1, d2, b1q, b2q, margin, u2, dec)
    return 1


def test_float_tier_is_certified_on_random_hits():
    rng = np.random.default_rng(5)
    decided = 0
    for _ in range(20000):
        b1q = rng.uniform(0.5, 5.0) * 2.0 ** -24
        b2q = rng.uniform(-0.6, 0.6) * 2.0 ** -24
        d1 = int(rng.integers(-(2 ** 28), 2 ** 28)) >> int(rng.integers(0, 28))
        d2 = int(rng.integers(-(2 ** 29), 2 ** 29)) >> int(rng.integers(0, 29))
        margin = 10.0 ** rng.uniform(-9, -4)
        u2 = int(rng.integers(0, 2 ** 32)) * 2.0 ** -32
        decided += check(d1, d2, b1q, b2q, margin, u2)
    assert decided > 15000  # the tier decides nearly every hit


def test_float_tier_near_every_threshold():
    """Uniforms within a few float ulps of exp(-dc), changes near 0, and the
    37 and 60 cut-offs: the tier must defer or agree with the exact test."""
    rng = np.random.default_rng(7)
    for _ in range(20000):
        b1q = rng.uniform(0.5, 5.0) * 2.0 ** -24
        b2q = rng.uniform(-0.6, 0.6) * 2.0 ** -24
        margin = 10.0 ** rng.uniform(-9, -5)
        target = rng.choice([rng.uniform(-1e-4, 1e-4), rng.uniform(0, 40), rng.uniform(36.9, 37.1),
                             rng.uniform(59.9, 60.1)])
        d2 = int(rng.integers(-(2 ** 20), 2 ** 20))
        d1 = int(round((target - b2q * d2) / b1q))
        dc = b1q * d1 + b2q * d2
        base = math.exp(-dc) if dc > 0 else 1.0 - abs(dc)
        u2 = min(max(base * (1.0 + rng.uniform(-3e-5, 3e-5)), 0.0), 1.0 - 2.0 ** -32)
        u2 = math.floor(u2 * 2.0 ** 32) * 2.0 ** -32  # a value the 32-bit draw can take
        check(d1, d2, b1q, b2q, margin, u2)


def test_wrong_band_is_caught():
    """With the float rounding left out of the band (Mf = margin only) some
    decision disagrees with the exact test: the checks above can fail."""
    rng = np.random.default_rng(11)
    bad = 0
    for _ in range(20000):
        b1q = rng.uniform(0.5, 5.0) * 2.0 ** -24
        b2q = rng.uniform(-0.6, 0.6) * 2.0 ** -24
        d2 = int(rng.integers(-(2 ** 29), 2 ** 29))
        d1 = int(rng.integers(-(2 ** 28), 2 ** 28))
        margin = 1e-12
        dc = b1q * d1 + b2q * d2
        if not 0 < dc < 30:
            continue
        u2 = math.floor(math.exp(-dc) * (1.0 + rng.uniform(-2e-7, 2e-7)) * 2.0 ** 32) * 2.0 ** -32
        p1, p2 = F(b1q) * F(d1), F(b2q) * F(d2)
        dcf, M = p1 + p2, F(margin)
        u = F(u2)
        e = np.exp2(-dcf * F(1.44269504))
        dec = 1 if u < e * (F(1) - M) else (0 if u >= e * (F(1) + M) else -1)  # no slack, no rounding term
        if dec >= 0 and exact(dc, u2) != dec:
            bad += 1
    assert bad > 0
