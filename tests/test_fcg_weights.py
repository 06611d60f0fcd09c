"""The flexible CG's sweep weights (tl_fcg_sweep_weights, no GPU work):
a fixed omega, or the Chebyshev step weights of [2 / a, 2] (tl_fcg_cheby).

Pinned by what the mathematics fixes, not by retyping the formula: the
weights are the reciprocals of the roots of the degree-m Chebyshev
polynomial mapped onto [lo, hi] (checked as roots: T_m vanishes there), so
the residual polynomial q(x) = prod_k (1 - w_k x) is T_m(sigma(x)) / T_m(sigma(0))
and its largest magnitude on [lo, hi] is 1 / T_m(sigma(0)) (sigma mapping
[lo, hi] onto [1, -1]), q(0) = 1, and the preconditioner polynomial
p(x) = (1 - q(x)) / x is positive on (0, 2], i.e. the sweeps stay a
symmetric positive definite preconditioner for D^-1 A's spectrum in (0, 2).
"""
import ctypes

import numpy as np
import pytest


def weights(hip, m, omega, ratio):
    w = (ctypes.c_double * m)()
    rc = hip.lib(2).tl_fcg_sweep_weights(ctypes.c_int(m), ctypes.c_double(omega), ctypes.c_double(ratio), w)
    assert rc == 0, hip.lib(2).tl_last_error()
    return np.array(w[:])


def test_fixed_weight(hip):
    assert np.array_equal(weights(hip, 4, 0.8, 0.0), [0.8] * 4)
    assert np.array_equal(weights(hip, 7, 0.65, 0.0), [0.65] * 7)


def partial_growth(w, x):
    """Largest |prod_{j <= k} (1 - w_j x)| over k and x (the iterate's
    growth) and over the suffixes (what a rounding error made before the
    last sweeps is multiplied by)."""
    pre = suf = 0.0
    q = np.ones_like(x)
    for wk in w:
        q = q * (1.0 - wk * x)
        pre = max(pre, np.abs(q).max())
    q = np.ones_like(x)
    for wk in w[::-1]:
        q = q * (1.0 - wk * x)
        suf = max(suf, np.abs(q).max())
    return pre, suf


@pytest.mark.parametrize("m,ratio", [(4, 30.0), (4, 4.0), (3, 100.0), (8, 30.0), (1, 2.0), (16, 3000.0), (12, 1000.0),
                                     (32, 13000.0)])
def test_chebyshev_weights(hip, m, ratio):
    w = weights(hip, m, 0.8, ratio)
    lo, hi = 2.0 / ratio, 2.0
    sigma = lambda x: (hi + lo - 2.0 * x) / (hi - lo)
    # 1 / w_k are the zeros of T_m(sigma(x)), each once
    assert np.all((1.0 / w > lo) & (1.0 / w < hi))
    assert np.all(np.diff(np.sort(1.0 / w)) > 0)
    tm = np.polynomial.chebyshev.Chebyshev.basis(m)
    assert np.max(np.abs(tm(sigma(1.0 / w)))) < 1e-12
    pre, suf = partial_growth(w, np.linspace(0.0, hi, 20001))
    if (m & (m - 1)) == 0:
        # the pairing order: both growths stay small, about linear in m
        # (decreasing order: the iterate within 1, but 5.5e7 on the
        # suffixes for m = 16)
        assert pre < 12 * m and suf < 48 * m, (pre, suf)
        if m >= 4:
            # each sweep k < m/2 of the first half pairs with the mirrored root
            th = 1.0 / w
            assert th[0] == th.max() and th[1] == th.min()
    else:
        assert np.all(np.diff(1.0 / w) < 0) and pre <= 1.0 + 1e-12
    x = np.linspace(lo, hi, 20001)
    q = np.prod(1.0 - np.outer(x, w), axis=1)
    bound = 1.0 / np.cosh(m * np.arccosh(sigma(0.0)))
    assert abs(np.max(np.abs(q)) - bound) <= 1e-9 * bound
    assert np.allclose(q, tm(sigma(x)) / tm(sigma(0.0)), rtol=0, atol=1e-12)
    xs = np.concatenate([np.logspace(-9, np.log10(lo), 200), x])
    qs = np.prod(1.0 - np.outer(xs, w), axis=1)
    assert np.all((1.0 - qs) / xs > 0)


def test_refusals(hip):
    w = (ctypes.c_double * 4)()
    L = hip.lib(2)
    for m, ratio in ((0, 0.0), (33, 0.0), (4, 1.0), (4, -3.0)):
        assert L.tl_fcg_sweep_weights(ctypes.c_int(m), ctypes.c_double(0.8), ctypes.c_double(ratio), w) == -1
        L.tl_clear_error()
