"""CPU pins of the time-blocked pass's weight algebra (the weights that
pair_block and pair_list in quest_amd/csrc/quest_kernels.hip form once per pair
and pass) against the oracle's per-pair coefficients, which follow the
reference's formulas (rotating_green_function.cpp:3-24, green_function.cpp:3-20,
green_function.h:32-55, lagrange_set.cpp:17-41).

The kernels rewrite the reference's coefficient on sample s - m,
    coef_m = d_b . G_m . d_a,
in units of c0 dt (tau = r / (c0 dt) = n + x, ir = 1 / tau) as
    coef_m = -2 K0 / (c0 dt^3) e^{-i theta tau} W_m      (rotating; lab: +2 K0 / (c0 dt^3) W_m)
    W_m = A L_m(x) + B L_m'(x) + C L_m''(x) / 2,
    A = g3h (ir + i theta) - g1 theta^2 / 2,  B = -g3h - i theta g1,  C = g1
    (lab: A = g3h ir, B = -g3h, C = g1),
    g3h = (d_a.d_b / 2 - 3/2 (d_a.rh)(d_b.rh)) ir^2,  g1 = (d_a.d_b - (d_a.rh)(d_b.rh)) ir,
with K0 = mu0 / (4 pi hbar) and theta = omega dt, and form W_m through the
monomial coefficients of the Lagrange basis on the centred nodes m - M/2 in
z = x - M/2: Gamma_j = A z^j + j B z^(j-1) + j(j-1)/2 C z^(j-2), and
W_m = E_m + O_m, W_{M-m} = E_m - O_m with E_m, O_m the sums of
c[j][m] Gamma_j over even and odd j.  This file restates that construction in
FP64 numpy, with the centred coefficients computed here as exact fractions, and
checks it against the oracle for random pairs, orders 1-6 and both frames,
so that a wrong sign, index, node offset or mirror fails on the CPU.  The GPU
tests check the kernels themselves against the oracle.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

C0, HBAR, MU0, OMEGA, DT = 299.792458, 0.65821193, 2.0133545e-4, 2278.9013, 1e-4


def centred_monomials(order):
    """c[j][m]: L_m(z) = sum_j c[j][m] z^j on the nodes m - order/2."""
    nodes = [Fraction(m) - Fraction(order, 2) for m in range(order + 1)]
    c = [[Fraction(0)] * (order + 1) for _ in range(order + 1)]
    for m, xm in enumerate(nodes):
        poly, den = [Fraction(1)], Fraction(1)
        for j, xj in enumerate(nodes):
            if j == m:
                continue
            nxt = [Fraction(0)] * (len(poly) + 1)
            for e, v in enumerate(poly):
                nxt[e + 1] += v
                nxt[e] -= xj * v
            poly, den = nxt, den * (xm - xj)
        for j in range(order + 1):
            c[j][m] = poly[j] / den
    return c


def blocked_weights(order, tau, da, db, rhat, theta, lab):
    """The kernels' W_m (FP64), m = 0..order."""
    n = math.floor(tau)
    x = tau - n
    ir = 1.0 / tau
    dr = rhat * tau
    ba, bb = float(da @ dr), float(db @ dr)
    aabh = 0.5 * float(da @ db)
    bq = ba * bb * ir * ir
    g3h = (-1.5 * bq + aabh) * ir * ir
    g1 = (2.0 * aabh - bq) * ir
    if lab:
        a, b, cc = g3h * ir, -g3h, g1
    else:
        rot = np.exp(-1j * theta * (n + 0.5)) * np.exp(-1j * theta * (x - 0.5))
        a = (g3h * (ir + 1j * theta) - g1 * theta * theta / 2) * rot
        b = (-g3h - 1j * theta * g1) * rot
        cc = g1 * rot
    z = x - order / 2
    gam = [a * z ** j + (j * b * z ** (j - 1) if j >= 1 else 0) + (j * (j - 1) / 2 * cc * z ** (j - 2) if j >= 2 else 0)
           for j in range(order + 1)]
    c = centred_monomials(order)
    w = np.zeros(order + 1, dtype=complex)
    for m in range(order + 1):
        if 2 * m >= order:
            continue
        e = sum(float(c[j][m]) * gam[j] for j in range(0, order + 1, 2))
        o = sum(float(c[j][m]) * gam[j] for j in range(1, order + 1, 2))
        w[m], w[order - m] = e + o, e - o
    if order % 2 == 0:
        w[order // 2] = sum(float(c[j][order // 2]) * gam[j] for j in range(0, order + 1, 2))
    return w


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_centred_coefficients_are_mirror_symmetric(order):
    """c[j][M - m] = (-1)^j c[j][m], the identity the mirror sums use, and
    the basis interpolates: L_m(node_k) = delta_mk."""
    c = centred_monomials(order)
    for m in range(order + 1):
        for j in range(order + 1):
            assert c[j][order - m] == (-1) ** j * c[j][m]
        for k in range(order + 1):
            zk = Fraction(k) - Fraction(order, 2)
            assert sum(c[j][m] * zk ** j for j in range(order + 1)) == (1 if k == m else 0)


@pytest.mark.parametrize("lab", [False, True], ids=["rotating", "lab"])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
def test_blocked_weights_match_oracle_coefficients(oracle_lib, order, lab):
    p = oracle_lib.params(C0, HBAR, MU0, OMEGA, DT, order, 22)
    theta = OMEGA * DT
    k0 = MU0 / (4 * math.pi * HBAR)
    scale = (1.0 if lab else -1.0) * 2.0 * k0 / (C0 * DT ** 3)
    rng = np.random.default_rng(100 + order + 10 * lab)
    worst = 0.0
    for _ in range(200):
        tau = float(rng.integers(order + 1, 300)) + float(rng.uniform(0.01, 0.99))
        rhat = rng.normal(size=3)
        rhat /= np.linalg.norm(rhat)
        da, db = rng.normal(size=3) * 5.29e-4, rng.normal(size=3) * 5.29e-4
        dr = rhat * tau * C0 * DT
        lw = oracle_lib.lagrange_weights(tau - math.floor(tau), order, DT)
        coefs = oracle_lib.lab_coefs if lab else oracle_lib.rotating_coefs
        want = coefs(p, dr, lw, db, da)
        got = scale * blocked_weights(order, tau, da, db, rhat, theta, lab)
        worst = max(worst, float(np.abs(got - want).max() / np.abs(want).max()))
    assert worst < 1e-11, worst


def test_three_product_sums_give_the_complex_sum():
    """The pass's step sums p1 = sum wr x, p2 = sum wi y, p3 = sum (wr + wi)(x + y)
    give re = p1 - p2 and im = p3 - p1 - p2 of sum w s."""
    rng = np.random.default_rng(7)
    w = rng.normal(size=64) + 1j * rng.normal(size=64)
    s = rng.normal(size=64) + 1j * rng.normal(size=64)
    p1 = np.sum(w.real * s.real)
    p2 = np.sum(w.imag * s.imag)
    p3 = np.sum((w.real + w.imag) * (s.real + s.imag))
    ref = np.sum(w * s)
    assert abs(complex(p1 - p2, p3 - p1 - p2) - ref) < 1e-12 * np.sum(np.abs(w * s))
