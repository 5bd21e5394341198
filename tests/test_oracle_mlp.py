"""The CPU restatement of the warm-start network (oracle/mlp.py), pinned by central
finite differences of its float64 path and by known BF16 roundings.  The network has no
counterpart in the reference; the GPU kernels are checked against this restatement in
tests/test_gpu_mlp.py."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    # ties to even at 1 + 2^-8 and 1 + 3 * 2^-8 (bfloat16 keeps 7 fraction bits)
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7 + 2 ** -9, 0.1], np.float32)
    assert np.array_equal(mlp.bf16(x), np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0 + 2 ** -7, 0.10009765625], np.float32))
    assert np.array_equal(mlp.bf16(mlp.bf16(x)), mlp.bf16(x))   # idempotent


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = mlp.init_params(3).astype(np.float64)
    n = 24
    feat = rng.standard_normal((n, mlp.IN))
    t = rng.standard_normal((n, 6))
    t[:, :3] /= np.linalg.norm(t[:, :3], axis=1, keepdims=True)
    t[:, 3:] /= np.linalg.norm(t[:, 3:], axis=1, keepdims=True)
    tgt = np.zeros((n, 8))
    tgt[:, 0:3], tgt[:, 3] = t[:, :3], 1.0
    tgt[:, 4:7], tgt[:, 7] = t[:, 3:], (rng.random(n) < 0.6).astype(float)
    loss, g = mlp.loss_and_grad(p, feat, tgt, rounded=False)
    assert 0.0 < loss < 4.0
    h = 1e-6
    idx = np.concatenate([rng.integers(lo, lo + n_, 8) for lo, n_ in
                          ((mlp.W1, mlp.H * mlp.IN), (mlp.W2, mlp.H * mlp.H), (mlp.W3, mlp.H * mlp.H),
                           (mlp.W4, 6 * mlp.H), (mlp.B1, mlp.H), (mlp.B4, 6))])
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp.loss_and_grad(q, feat, tgt, rounded=False)
        q[i] -= 2 * h
        lm, _ = mlp.loss_and_grad(q, feat, tgt, rounded=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # unused outputs 6..15 get no gradient
    assert not np.any(g[mlp.W4 + 6 * mlp.H:mlp.B4]) and not np.any(g[mlp.B4 + 6:])


def test_prediction_is_normalised():
    rng = np.random.default_rng(5)
    out = mlp.predict(mlp.init_params(1), rng.standard_normal((10, mlp.IN)))
    assert np.allclose(np.linalg.norm(out[:, :3], axis=1), 1.0) and np.allclose(np.linalg.norm(out[:, 3:], axis=1), 1.0)
