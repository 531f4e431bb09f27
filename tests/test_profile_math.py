"""The mathematics the range-profile kernels rely on, checked on the CPU in
FP64 (NumPy/SciPy), independently of the kernels:

* the cubic coefficients k_prof_coef stores, (q, c1, c2, c3) with
  c1 = -p/3 - q/2 + r - s/6, c2 = (p + r)/2 - q, c3 = (s - p)/6 + (q - r)/2,
  pass through the four table points at f = -1, 0, 1, 2;
* interpolating a row's band-centred range profile from samples h =
  c / (64 (F-1) df) apart with that cubic reproduces the exact sum to the
  stated bound (Lagrange remainder of the fastest tone, 0.0234 (pi/64)^4
  = 1.4e-7 of sum_j |s_j|);
* the grouped-series backward's truncations: 12 Chebyshev terms of
  exp(+j y xi) (Jacobi-Anger, with scipy's Bessel J_k) stay below 1e-8 for
  |y| <= 2, |xi| <= 1, and the series k_bt_tables evaluates for J_k agrees
  with scipy;
* the position term's table: the power-basis coefficients k_bt_tables forms
  from the 12 Chebyshev terms (w_k = sum_m (-1)^((m-k)/2) eps_m J_m(y) t_mk)
  give exp(-j y xi) within 1e-8 in FP64 and 1e-7 once stored as float32, as
  the 16-term power series y^k / k! does, and reduce to it as y -> 0.
"""
import math

import numpy as np
from scipy.special import jv

C0 = 2.99792458e8


def cubic_coeffs(p, q, r, s):
    return q, -p / 3 - q / 2 + r - s / 6, (p + r) / 2 - q, (s - p) / 6 + (q - r) / 2


def test_cubic_coefficients_hit_the_four_points():
    rng = np.random.default_rng(0)
    p, q, r, s = rng.normal(size=4) + 1j * rng.normal(size=4)
    c0, c1, c2, c3 = cubic_coeffs(p, q, r, s)
    P = lambda f: c0 + c1 * f + c2 * f ** 2 + c3 * f ** 3
    for f, want in ((-1, p), (0, q), (1, r), (2, s)):
        assert abs(P(f) - want) <= 1e-14 * (abs(p) + abs(q) + abs(r) + abs(s))


def test_profile_interpolation_error_bound():
    """One row, F = 256 samples 77-81 GHz, random samples; distances over
    0.55-0.85 m.  Exact P(d) against table + cubic + carrier, as the kernels
    compose them (table Q at d_lo + k h, interval m, fraction f)."""
    rng = np.random.default_rng(1)
    F = 256
    f0, df = 77e9, 4e9 / (F - 1)
    s = rng.normal(size=F) + 1j * rng.normal(size=F)
    tc = (F - 1) / 2
    fc = f0 + tc * df
    h = min(C0 / (64 * (F - 1) * df), 0.5 * C0 / fc)
    d_lo = 0.55 - h
    M = int(math.ceil(0.30 / h)) + 6
    dk = d_lo + h * np.arange(M)
    j = np.arange(F)
    Q = np.exp(-2j * np.pi * np.outer(dk, j - tc) * df / C0) @ s
    d = rng.uniform(0.56, 0.84, size=4000)
    exact = np.exp(-2j * np.pi * np.outer(d, f0 + j * df) / C0) @ s
    u = (d - d_lo) / h
    m = np.floor(u).astype(int)
    f = u - m
    c0, c1, c2, c3 = cubic_coeffs(Q[m - 1], Q[m], Q[m + 1], Q[m + 2])
    approx = (c0 + f * (c1 + f * (c2 + f * c3))) * np.exp(-2j * np.pi * fc * d / C0)
    err = np.abs(approx - exact) / np.sum(np.abs(s))
    assert err.max() <= 0.0234 * (np.pi / 64) ** 4
    assert np.sqrt(np.mean(err ** 2)) <= 5e-8


def test_series_truncations():
    y = np.linspace(-2.0, 2.0, 41)[:, None]
    xi = np.linspace(-1.0, 1.0, 41)[None, :]
    # Jacobi-Anger: exp(+j y cos t) = J0 + 2 sum_k j^k J_k(y) cos(k t), 12 terms
    T = [np.ones_like(xi), xi]
    for k in range(2, 12):
        T.append(2 * xi * T[-1] - T[-2])
    ch = sum((1 if k == 0 else 2) * (1j ** k) * jv(k, y) * T[k] for k in range(12))
    assert np.max(np.abs(ch - np.exp(1j * y * xi))) <= 1e-8


def test_bessel_series_of_the_tables():
    """J_k(y) from 20 terms of sum_m (-1)^m (y/2)^(2m+k) / (m! (m+k)!), |y| <= 2,
    k < 12 (k_bt_tables), against scipy."""
    for y in np.linspace(-2.0, 2.0, 17):
        h = y / 2
        for k in range(12):
            term = h ** k / math.factorial(k)
            tot = 0.0
            for mm in range(20):
                tot += term
                term *= -h * h / ((mm + 1) * (mm + 1 + k))
            assert abs(tot - jv(k, y)) <= 1e-15


def power_basis_table(y, K=12):
    """The position term's table as k_bt_tables forms it: eps_m J_m(y) from
    20 terms of the Bessel series, then the power-basis coefficients of
    sum_m eps_m (-j)^m J_m(y) T_m(xi) without the (-j)^k factor."""
    h = y / 2
    ej = []
    for k in range(K):
        term, tot = h ** k / math.factorial(k), 0.0
        for mm in range(20):
            tot += term
            term *= -h * h / ((mm + 1) * (mm + 1 + k))
        ej.append((2.0 if k else 1.0) * tot)
    # T_m coefficients from numpy's Chebyshev-to-power conversion
    w = np.zeros(K)
    for m in range(K):
        t = np.polynomial.chebyshev.cheb2poly([0] * m + [1])
        for k in range(m % 2, m + 1, 2):
            w[k] += (-1) ** ((m - k) // 2) * ej[m] * t[k]
    return w


def test_position_table_power_basis():
    xi = np.linspace(-1.0, 1.0, 201)
    worst64 = worst32 = 0.0
    for y in np.linspace(-2.0, 2.0, 41):
        w = power_basis_table(y)
        exact = np.exp(-1j * y * xi)
        for ww, f32 in ((w, False), (w.astype(np.float32).astype(np.float64), True)):
            P = sum((-1j) ** k * ww[k] * xi ** k for k in range(len(ww)))
            err = np.max(np.abs(P - exact))
            if f32:
                worst32 = max(worst32, err)
            else:
                worst64 = max(worst64, err)
    assert worst64 <= 1e-8
    assert worst32 <= 1e-7
    # small y: the coefficients approach the power series y^k / k!
    y = 1e-3
    w = power_basis_table(y)
    taylor = np.array([y ** k / math.factorial(k) for k in range(12)])
    assert np.max(np.abs(w - taylor)) <= 1e-15
