"""The network's 3x3 convolutions with the input tile staged in LDS against
the same convolutions reading through L2 once per tap (tl_learn_conv_staging).

Both run the same MFMA sequence for every output, so the forward pass is
bitwise equal.  The gradients pass through the weight-gradient kernel's
atomic sums, whose order varies from run to run, so they are compared per
layer at 1e-5 relative L2.  Shapes cover the staged tile's edges: widths
below and not a multiple of 64, heights not a multiple of the tile's 4, 2 or
1 rows, batches, and every layer width (8 to 192 input channels).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(2, 16, 16), (1, 40, 72), (3, 24, 136), (1, 136, 200)]


@pytest.fixture(scope="module")
def learner(hip):
    L = hip.Learner(seed=21)
    yield L
    hip.lib(2).tl_learn_conv_staging(1)
    L.close()


def _planes(shape, seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=shape).astype(np.float32)
    f = rng.uniform(-1, 1, size=shape).astype(np.float32)
    t = (0.3 * rng.normal(size=shape)).astype(np.float32)
    return r, f, t


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_bitwise(hip, learner, shape):
    print(f"unet staged forward {shape}", flush=True)
    r, f, _ = _planes(shape, 1)
    L = hip.lib(2)
    L.tl_learn_conv_staging(1)
    z1 = learner.apply(r, f)
    L.tl_learn_conv_staging(0)
    z0 = learner.apply(r, f)
    L.tl_learn_conv_staging(1)
    assert np.isfinite(z1).all() and np.abs(z1).max() > 0
    assert np.array_equal(z1.view(np.uint32), z0.view(np.uint32))


@pytest.mark.parametrize("shape", SHAPES[:3])
def test_gradients_match(hip, learner, shape):
    print(f"unet staged gradients {shape}", flush=True)
    r, f, t = _planes(shape, 2)
    rng = np.random.default_rng(3)
    kx, ky = rng.uniform(0.05, 20.0, size=(2,) + shape).astype(np.float32)
    L = hip.lib(2)
    L.tl_learn_conv_staging(1)
    l1, g1 = learner.loss_grad(r, f, t, kx, ky)
    L.tl_learn_conv_staging(0)
    l0, g0 = learner.loss_grad(r, f, t, kx, ky)
    L.tl_learn_conv_staging(1)
    assert abs(l1 - l0) <= 1e-6 * abs(l0)
    import unet_oracle as U
    k = 0
    for cin, cout, _ in U.SHAPES:
        n = cout * 9 * cin + cout
        a, b = g1[k:k + n].astype(np.float64), g0[k:k + n].astype(np.float64)
        k += n
        nb = np.linalg.norm(b)
        assert np.linalg.norm(a - b) <= 1e-5 * nb + 1e-30, (cin, cout)
