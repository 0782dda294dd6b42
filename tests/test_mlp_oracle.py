"""The numpy oracle of the perceptron decoder (oracle/mlp.py) against finite
differences of its own loss (float64, no BF16 rounding) on a tiny network,
and its BF16 rounding against known values."""
import numpy as np

from oracle import mlp as om


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -3.14159265, 0.0], np.float32)
    got = om.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even (1 + 2^-6) wins
    assert got.tolist() == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -3.140625, 0.0]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = (5, 128, 128, 128)
    kin, shapes, offs, n = om.layout(*dims)
    p = rng.normal(0, 0.2, n)
    X = np.zeros((16, kin))
    X[:, :5] = rng.integers(0, 2, (16, 5))
    y = rng.integers(0, 2, 16).astype(float)
    loss, g = om.loss_and_grads(p, X, y, dims, emulate=False)
    idx = [offs[0] + 3, offs[0] + kin + 1, offs[1] + 7, offs[2] + 129, offs[3] + 5,
           offs[4] + 1000, offs[5] + 2, offs[6] + 17, offs[7]]
    for i in idx:
        e = np.zeros(n)
        e[i] = 1e-6
        lp, _ = om.loss_and_grads(p + e, X, y, dims, emulate=False)
        lm, _ = om.loss_and_grads(p - e, X, y, dims, emulate=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])
    # padding input columns see zero inputs: zero gradient
    assert np.all(g[offs[0]:offs[1]].reshape(128, kin)[:, 5:] == 0)


def test_adam_first_step_moves_by_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 0.0])
    q, m, v = om.adam(p, np.zeros(3), np.zeros(3), g, 1, 1e-2)
    # bias-corrected first step is lr * sign(g) (for |g| >> eps)
    assert np.allclose(q[:2], p[:2] - 1e-2 * np.sign(g[:2]), atol=1e-8)
    assert q[2] == p[2]
