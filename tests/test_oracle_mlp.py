"""CPU checks of the warm-start network's oracle (numpy FP64): GELU identities,
forward known answers and central finite differences of the loss gradient."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_gelu_identities():
    z = np.linspace(-6, 6, 241)
    # gelu(z) - gelu(-z) = z, gelu(0) = 0, gelu'(0) = 1/2, tails
    np.testing.assert_allclose(mo.gelu(z) - mo.gelu(-z), z, atol=1e-14)
    assert mo.gelu(0.0) == 0.0 and mo.gelu_grad(0.0) == 0.5
    assert abs(mo.gelu(8.0) - 8.0) < 1e-12 and abs(mo.gelu(-8.0)) < 1e-12
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h), atol=1e-8)


def test_param_layout_size():
    assert mo.N_PARAMS == 35459


def test_forward_known_answers():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(7, 16))
    p = np.zeros(mo.N_PARAMS)
    p[mo.O_B4:] = [1.0, -2.0, 0.5]
    np.testing.assert_allclose(mo.forward(p, X), np.tile([1.0, -2.0, 0.5], (7, 1)))
    # one path: x0 -> unit 0 of every layer -> output 0 (others zero)
    p = np.zeros(mo.N_PARAMS)
    p[mo.O_W1 + 0] = 1.0
    p[mo.O_W2 + 0] = 1.0
    p[mo.O_W3 + 0] = 1.0
    p[mo.O_W4 + 0] = 1.0
    y = mo.forward(p, X)
    np.testing.assert_allclose(y[:, 0], mo.gelu(mo.gelu(mo.gelu(X[:, 0]))))
    assert np.all(y[:, 1:] == 0)
    # the padding column of the feature row is ignored
    X2 = X.copy()
    X2[:, 15] = 99.0
    p = np.random.default_rng(1).normal(size=mo.N_PARAMS) * 0.1
    np.testing.assert_array_equal(mo.forward(p, X), mo.forward(p, X2))


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_by_central_differences(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=mo.N_PARAMS) * 0.2
    X = rng.normal(size=(5, 16))
    T = rng.normal(size=(5, 3))
    loss, g = mo.loss_grad(p, X, T)
    idx = rng.choice(mo.N_PARAMS, size=60, replace=False)
    idx = np.concatenate([idx, [mo.O_B1, mo.O_B2 + 3, mo.O_B3 + 5, mo.O_B4, mo.O_B4 + 2, mo.O_W1 + 14]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss_grad(p + e, X, T)[0] - mo.loss_grad(p - e, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert loss > 0


def test_features_definition_small():
    # 2 vertices, one block row each plus coupling; hand-computable rows
    ptr = np.array([0, 2, 4])
    col = np.array([0, 1, 0, 1])
    blocks = np.zeros((4, 3, 3))
    blocks[0] = np.diag([2.0, 4.0, 6.0])
    blocks[0][0, 1] = blocks[0][1, 0] = 0.5
    blocks[3] = np.diag([1.0, 1.0, 1.0])
    res = np.array([4.0, 0, 0, 0, 3.0, 0])
    dprev = np.zeros(6)
    F, s = mo.features(res, ptr, col, blocks, dprev)
    # d0 = 4, y0 = (1, 0, 0); d1 = 1, y1 = (0, 3, 0); s = sqrt((1 + 9) / 2)
    assert abs(s - np.sqrt(5.0)) < 1e-15
    np.testing.assert_allclose(F[0, 0:3], np.array([1, 0, 0]) / s)
    np.testing.assert_allclose(F[0, 3:9], [0.5, 1.0, 1.5, 0.125, 0.0, 0.0])
    np.testing.assert_allclose(F[0, 12:15], np.array([0, 3, 0]) / s)  # neighbour 1
    np.testing.assert_allclose(F[1, 12:15], np.array([1, 0, 0]) / s)  # neighbour 0
    assert np.all(F[:, 15] == 0)
