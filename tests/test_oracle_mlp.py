"""The closure network's CPU oracle (oracle/mlp_ref.py), pinned by finite
differences, bf16 rounding of known values, and a linear layer's closed form."""
import numpy as np

from oracle import mlp_ref


def test_bf16_rounding_known_values():
    # bf16 keeps 7 fraction bits: 1 + 2^-7 is exact; 1 + 2^-8 is a tie that goes
    # down to the even 1.0; 1 + 3 * 2^-8 is a tie that goes up to the even 1 + 2^-6;
    # 1 + 3 * 2^-9 is above the midpoint and goes up to 1 + 2^-7.
    x = np.array([1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5,
                  0.1, 0.0], dtype=np.float32)
    want = [1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.10009765625, 0.0]
    np.testing.assert_array_equal(mlp_ref.bf16(x), want)


def test_exact_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    widths = [5, 7, 6, 2]
    p = mlp_ref.glorot_params(widths, rng)
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 2))
    _, g = mlp_ref.loss_and_grad(p, widths, x, t, exact=True)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = eps
        lp, _ = mlp_ref.loss_and_grad(p + e, widths, x, t)
        lm, _ = mlp_ref.loss_and_grad(p - e, widths, x, t)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.abs(fd - g).max() < 1e-8 * max(1.0, np.abs(g).max())


def test_linear_layer_closed_form():
    """One layer without activation: y = x W^T + b and dL/dW = 2/n (y - t)^T x."""
    rng = np.random.default_rng(4)
    x, t = rng.normal(size=(8, 3)), rng.normal(size=(8, 2))
    w, b = rng.normal(size=(2, 3)), rng.normal(size=2)
    p = np.concatenate([w.reshape(-1), b])
    y, _ = mlp_ref.forward(p, [3, 2], x)
    np.testing.assert_allclose(y, x @ w.T + b, rtol=1e-14)
    loss, g = mlp_ref.loss_and_grad(p, [3, 2], x, t)
    d = 2.0 * (y - t) / y.size
    np.testing.assert_allclose(g[:6], (d.T @ x).reshape(-1), rtol=1e-13)
    np.testing.assert_allclose(g[6:], d.sum(0), rtol=1e-13)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-15


def test_bf16_mode_close_to_exact():
    rng = np.random.default_rng(5)
    widths = [6, 32, 32, 3]
    p = mlp_ref.glorot_params(widths, rng)
    x, t = rng.normal(size=(64, 6)), rng.normal(size=(64, 3))
    ye, _ = mlp_ref.forward(p, widths, x, exact=True)
    yb, _ = mlp_ref.forward(p, widths, x, exact=False)
    assert np.abs(ye - yb).max() < 0.05 * np.abs(ye).max()
    _, ge = mlp_ref.loss_and_grad(p, widths, x, t, exact=True)
    _, gb = mlp_ref.loss_and_grad(p, widths, x, t, exact=False)
    assert np.linalg.norm(ge - gb) < 0.05 * np.linalg.norm(ge)


def test_log2_scaled_form_equals_the_network():
    """The fused kernel's reformulation (log2-scaled activations and biases, layer
    0's bias as a high + low pair) is the same function: exact mode equals the
    plain forward; with bf16 rounding both stay at bf16-level distance."""
    rng = np.random.default_rng(11)
    for widths in ([40, 256, 256, 256, 15], [13, 64, 64, 6], [80, 512, 512, 30]):
        p = mlp_ref.glorot_params(widths, rng)
        p[-widths[-1]:] = rng.normal(size=widths[-1])  # output biases of order 1
        x = rng.normal(size=(64, widths[0]))
        y, _ = mlp_ref.forward(p, widths, x, exact=True)
        y2 = mlp_ref.forward_log2_scaled(p, widths, x, exact=True)
        np.testing.assert_allclose(y2, y, rtol=0, atol=1e-12 * np.abs(y).max())
        yb, _ = mlp_ref.forward(p, widths, x, exact=False)
        y2b = mlp_ref.forward_log2_scaled(p, widths, x, exact=False)
        scale = np.abs(y).max()
        assert np.abs(y2b - y).max() < 3e-2 * scale
        assert np.abs(y2b - yb).max() < 3e-2 * scale
        # a dropped log2(e) or bias would be far off
        assert np.abs(mlp_ref.forward_log2_scaled(p * 1.01, widths, x) - y).max() > 1e-3 * scale
