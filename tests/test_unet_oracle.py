"""The CPU oracle of the learned preconditioner's network, pinned by
mathematics rather than by restating it: the backward pass against central
finite differences of the loss (FP64 mode, no rounding), the BF16 rounding
against known values, and the convolution against a direct loop."""
import numpy as np
import pytest

import unet_oracle as U


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 3.0e38, 1.0e-40], dtype=np.float64)
    got = U.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -6
    assert got[3] == -2.5 and got[4] == 0.0
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2.0 ** -8
    # every result is representable with 8 significant bits
    m, _ = np.frexp(got[np.isfinite(got) & (got != 0)])
    assert np.all(np.abs(m * 256 - np.round(m * 256)) == 0)


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(4, 9, 3))
    y = U.conv(x, w)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for yy in range(5):
            for xx in range(6):
                for t in range(9):
                    sy, sx = yy + t // 3 - 1, xx + t % 3 - 1
                    if 0 <= sy < 5 and 0 <= sx < 6:
                        ref[n, yy, xx] += w[:, t, :] @ x[n, sy, sx]
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    # the adjoint identity <conv(x), g> = <x, conv_bwd_data(g)>
    g = rng.normal(size=y.shape)
    assert np.isclose(np.sum(y * g), np.sum(x * U.conv_bwd_data(g, w)), rtol=1e-12)
    # and the weight gradient is the derivative of <conv(x, w), g> in w
    dw = U.conv_bwd_weight(g, x)
    e = np.zeros_like(w)
    e[2, 4, 1] = 1.0
    assert np.isclose(np.sum(U.conv(x, e) * g), dw[2, 4, 1], rtol=1e-12)


def test_a_apply_is_the_symmetric_operator():
    """A = I + sum of face fluxes: symmetric, A 1 = 1, and equal to the
    assembled matrix of the same face coefficients."""
    rng = np.random.default_rng(5)
    kx, ky = rng.uniform(0.1, 3.0, size=(2, 1, 4, 5))
    n = 20
    a = np.zeros((n, n))
    for i in range(n):
        e = np.zeros((1, 4, 5))
        e.flat[i] = 1.0
        a[:, i] = U.a_apply(e, kx, ky).ravel()
    assert np.allclose(a, a.T)
    assert np.allclose(U.a_apply(np.ones((1, 4, 5)), kx, ky), 1.0)
    # the (x, y) = (2, 1) row: diagonal 1 + its four face coefficients
    i = 1 * 5 + 2
    assert np.isclose(a[i, i], 1 + kx[0, 1, 2] + kx[0, 1, 3] + ky[0, 1, 2] + ky[0, 2, 2])
    assert np.isclose(a[i, i + 1], -kx[0, 1, 3]) and np.isclose(a[i, i - 5], -ky[0, 1, 2])


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1, 4, 6, 2))
    g = rng.normal(size=(1, 2, 3, 2))
    assert np.isclose(np.sum(U.pool(x) * g), np.sum(x * U.pool_bwd(g)))
    h = rng.normal(size=(1, 4, 6, 2))
    assert np.isclose(np.sum(U.up(g) * h), np.sum(g * U.up_bwd(h)))


def test_smoothing_operator_and_its_transpose():
    """smooth() is the error propagation of weighted Jacobi sweeps on the
    assembled operator (M = (I - omega D^-1 A)^m with D = diag(A)) and
    smooth_t() is its transpose: <M v, g> = <v, M^T g>."""
    rng = np.random.default_rng(8)
    kx, ky = rng.uniform(0.1, 4.0, size=(2, 1, 5, 6))
    n = 30
    A = np.zeros((n, n))
    for i in range(n):
        e = np.zeros((1, 5, 6))
        e.flat[i] = 1.0
        A[:, i] = U.a_apply(e, kx, ky).ravel()
    D = np.diag(np.diag(A))
    M = np.linalg.matrix_power(np.eye(n) - 0.7 * np.linalg.solve(D, A), 3)
    v = rng.normal(size=(1, 5, 6))
    g = rng.normal(size=(1, 5, 6))
    assert np.allclose(U.smooth(v, kx, ky, 3, 0.7).ravel(), M @ v.ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(U.smooth_t(g, kx, ky, 3, 0.7).ravel(), M.T @ g.ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(U.a_diag(kx, ky).ravel(), np.diag(A))


@pytest.mark.parametrize("seed,anorm,sweeps", [(1, False, 0), (2, False, 0), (3, True, 0), (4, True, 3)])
def test_backward_matches_finite_differences(seed, anorm, sweeps):
    rng = np.random.default_rng(seed)
    params = U.init_params(seed).astype(np.float64)
    # positive biases keep the deep 1x1 levels of an 8x8 input alive
    k = 0
    for cin, cout, _ in U.SHAPES:
        k += cout * 9 * cin
        params[k:k + cout] = 0.05
        k += cout
    ns = 2 if sweeps else 1  # two samples pin the per-sample normalisation
    r = rng.normal(size=(ns, 8, 8))
    f = rng.uniform(-1, 1, size=(ns, 8, 8))
    t = rng.normal(size=(ns, 8, 8)) * np.arange(1, ns + 1)[:, None, None]
    kx = ky = None
    if anorm:
        kx, ky = rng.uniform(0.1, 5.0, size=(2, ns, 8, 8))
    loss, grad = U.loss_grad(params, r, f, t, bf=False, kx=kx, ky=ky, sweeps=sweeps)
    # parameters spread over all layers, away from unused input channels
    picks, k = [], 0
    for i, (cin, cout, _) in enumerate(U.SHAPES):
        nw = cout * 9 * cin
        for _ in range(2):
            co, tap, ci = rng.integers(cout), rng.integers(9), rng.integers(2 if i == 0 else cin)
            picks.append(k + (co * 9 + tap) * cin + ci)
        picks.append(k + nw + rng.integers(cout))
        k += nw + cout
    checked = 0

    def fd(j, h):
        p1, p2 = params.copy(), params.copy()
        p1[j] += h
        p2[j] -= h
        return (U.loss_grad(p1, r, f, t, bf=False, kx=kx, ky=ky, sweeps=sweeps)[0]
                - U.loss_grad(p2, r, f, t, bf=False, kx=kx, ky=ky, sweeps=sweeps)[0]) / (2 * h)

    for j in picks:
        a, b = fd(j, 1e-6), fd(j, 2e-7)
        if abs(a - b) > 1e-4 * max(abs(a), abs(b), 1e-12):
            continue  # a ReLU kink lies within the step
        if abs(a) < 1e-9 and abs(grad[j]) < 1e-9:
            continue
        assert abs(a - grad[j]) <= 1e-5 * max(abs(a), abs(grad[j])) + 1e-10, (j, a, grad[j])
        checked += 1
    assert checked >= len(picks) // 2
