"""The closure network on the GPU (BF16 MFMA GEMMs) against the numpy oracle.

Exactly representable operands make the MFMA products exact, so a wrong operand
lane layout, a transposed operand or a dropped K tile fails bit-exactly; random
operands are compared with the oracle's bf16 emulation."""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def exact_linear(rng, k, n, b_rows):
    """A one-layer net whose inputs, weights and every partial sum are exact in
    bf16 / fp32: entries are multiples of 1/4 in [-2, 2]."""
    x = rng.integers(-8, 9, size=(b_rows, k)) / 4.0
    w = rng.integers(-8, 9, size=(n, k)) / 4.0
    b = rng.integers(-8, 9, size=n) / 4.0
    return x, w, b


@pytest.mark.parametrize("k,n,rows", [(16, 32, 32), (80, 70, 130), (33, 1, 5), (200, 129, 257)])
def test_gemm_forward_bit_exact(gpu, k, n, rows):
    ce = gpu
    rng = np.random.default_rng(k * 1000 + n)
    x, w, b = exact_linear(rng, k, n, rows)
    net = ce.MLP([k, n], seed=1)
    net.set_params(np.concatenate([w.reshape(-1), b]))
    y = net.forward(x)
    np.testing.assert_array_equal(y, (x @ w.T + b).astype(np.float32))


def test_gemm_backward_bit_exact(gpu):
    """dW = dY^T X and db = sum dY with dY = 2 (Y - T) / (rows * outputs) exact."""
    ce = gpu
    rng = np.random.default_rng(11)
    k, n, rows = 48, 16, 64  # rows * n = 1024: the 2/1024 scale is a power of two
    x, w, b = exact_linear(rng, k, n, rows)
    y = x @ w.T + b
    t = y - rng.integers(-16, 17, size=y.shape) / 4.0  # residuals exact in bf16
    net = ce.MLP([k, n], seed=2)
    net.set_params(np.concatenate([w.reshape(-1), b]))
    loss, g = net.gradients(x, t)
    d = 2.0 * (y - t) / y.size
    np.testing.assert_array_equal(g[:k * n], (d.T @ x).reshape(-1).astype(np.float32))
    np.testing.assert_array_equal(g[k * n:], d.sum(0).astype(np.float32))
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-6)


@pytest.mark.parametrize("widths,rows", [([7, 64, 50, 3], 300), ([13, 256, 256, 256, 6], 1000)])
def test_forward_and_gradients_vs_oracle(gpu, widths, rows):
    ce = gpu
    rng = np.random.default_rng(len(widths))
    net = ce.MLP(widths, seed=3)
    p = mlp_ref.glorot_params(widths, rng).astype(np.float32)
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)
    x = rng.normal(size=(rows, widths[0])).astype(np.float32)
    t = rng.normal(size=(rows, widths[-1])).astype(np.float32)
    net.set_fused(False)  # the training path's layer-by-layer GEMMs (the fused kernel: below)
    y = net.forward(x)
    yb, _ = mlp_ref.forward(p, widths, x, exact=False)
    # Same bf16 roundings as the oracle; differences are FP32 accumulation order and
    # the odd activation that rounds to the other bf16 neighbour.
    assert np.abs(y - yb).max() < 1e-2 * np.abs(yb).max()
    assert np.sqrt(np.mean((y - yb) ** 2)) < 2e-3 * np.sqrt(np.mean(yb ** 2))
    loss, g = net.gradients(x, t)
    lb, gb = mlp_ref.loss_and_grad(p, widths, x, t, exact=False)
    assert loss == pytest.approx(lb, rel=1e-3)
    assert np.linalg.norm(g - gb) < 1e-2 * np.linalg.norm(gb)
    # and the mixed-precision gradient is a usable descent direction of the exact loss
    _, ge = mlp_ref.loss_and_grad(p, widths, x, t, exact=True)
    assert g @ ge > 0.95 * np.linalg.norm(g) * np.linalg.norm(ge)


def test_adam_step_matches_formula(gpu):
    ce = gpu
    widths = [5, 32, 4]
    rng = np.random.default_rng(8)
    net = ce.MLP(widths, seed=4)
    p0 = net.params().astype(np.float64)
    x, t = rng.normal(size=(64, 5)), rng.normal(size=(64, 4))
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0.copy()
    for step in range(1, 4):
        _, g = net.gradients(x, t)
        net.adam_step(1e-3)
        g = g.astype(np.float64)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
        np.testing.assert_allclose(net.params(), p, rtol=1e-5, atol=1e-7)
        # the bf16 weight copy follows the master parameters
        yb, _ = mlp_ref.forward(net.params(), widths, x, exact=False)
        assert np.abs(net.forward(x) - yb).max() < 1e-2 * np.abs(yb).max()


def test_train_epoch_reduces_loss(gpu):
    ce = gpu
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(8192, 4)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], x[:, 3] ** 2], axis=1).astype(np.float32)
    net = ce.MLP([4, 64, 64, 2], seed=5)
    net.set_dataset(x, t)
    first = net.train_epoch(256, 3e-3, shuffle_seed=0)
    for e in range(1, 15):
        last = net.train_epoch(256, 3e-3, shuffle_seed=e)
    assert last < 0.25 * first  # fp32 torch on the same setup: 0.168 -> 0.013
    y = net.forward(x[:512])
    assert np.mean((y - t[:512]) ** 2) < 2 * last + 1e-4


def test_one_batch_epoch_equals_gradient_step(gpu):
    """An epoch with a single whole-set batch is one gradient + Adam step."""
    ce = gpu
    rng = np.random.default_rng(10)
    widths = [3, 32, 2]
    x = rng.normal(size=(128, 3)).astype(np.float32)
    t = rng.normal(size=(128, 2)).astype(np.float32)
    a, b = ce.MLP(widths, seed=6), ce.MLP(widths, seed=6)
    np.testing.assert_array_equal(a.params(), b.params())
    a.set_dataset(x, t)
    la = a.train_epoch(128, 1e-2, shuffle_seed=1)
    lb, _ = b.gradients(x, t)
    b.adam_step(1e-2)
    assert la == pytest.approx(lb, rel=1e-5)
    # Rows are summed in another order in the gradients; Adam's first step is
    # lr * g / (|g| + eps), so only components with |g| near eps can move apart.
    d = np.abs(a.params() - b.params())
    assert np.mean(d < 1e-6) > 0.99 and d.max() < 2e-2


def test_mlp_argument_errors(gpu):
    ce = gpu
    with pytest.raises(ValueError):
        ce.MLP([4])
    with pytest.raises(ValueError):
        ce.MLP([4, 0, 2])
    net = ce.MLP([4, 8, 2])
    with pytest.raises(ValueError):
        net.forward(np.zeros((3, 5), np.float32))
    with pytest.raises(RuntimeError):
        net.train_epoch(16, 1e-3)  # no training set


@pytest.mark.parametrize("widths,rows", [([40, 256, 256, 256, 15], 1000),
                                         ([80, 512, 512, 512, 512, 30], 300),
                                         ([50, 512, 512, 512, 9], 200),
                                         ([13, 256, 256, 6], 257)])
def test_fused_forward_vs_oracle(gpu, widths, rows):
    """The fused inference kernel against the oracle's statement of its arithmetic
    (log2-scaled bf16 activations, layer 0's bias as a bf16 pair): equal up to the
    FP32 summation order and the 1-ulp exp2 / reciprocal.  Those differences are
    ~1e-8 of the output scale, except where they move a hidden activation across a
    bf16 rounding boundary (one step, 2^-8 of it; about one activation in 10^4),
    which shifts the outputs downstream of it by up to ~1e-3 of the scale (the
    oracle with float32 instead of exact sums shows the same: for the 5-layer
    512-wide net the 99th percentile of that shift is 3.2e-4 of the scale, 1.5e-4
    for the layer-by-layer form).  The bar: median within 1e-6 (a systematic error
    would move it), 99th percentile within 1e-3, largest within 1e-2 of the scale.
    Against the layer-by-layer path (SiLU rounded to bf16 unscaled): bf16 rounding
    of the activations."""
    ce = gpu
    rng = np.random.default_rng(rows)
    net = ce.MLP(widths, seed=9)
    p = mlp_ref.glorot_params(widths, rng).astype(np.float32)
    net.set_params(p)
    x = rng.normal(size=(rows, widths[0])).astype(np.float32)
    assert net.set_fused(True)
    y_fused = net.forward(x)
    assert not net.set_fused(False)
    y_layers = net.forward(x)
    scale = np.abs(y_layers).max()
    yf_ref = mlp_ref.forward_log2_scaled(p, widths, x, exact=False)
    d = np.abs(y_fused - yf_ref)
    yb, _ = mlp_ref.forward(p, widths, x, exact=False)
    d_layers = np.abs(y_layers - yb)
    assert d_layers.max() < 1e-2 * scale
    assert d.max() < 1e-2 * scale
    assert np.quantile(d, 0.99) < 1e-3 * scale
    assert np.median(d) < 1e-6 * scale
    assert np.abs(y_fused - y_layers).max() < 2e-2 * scale
    # a row's result depends on its own inputs only: a call over the first rows
    # (a partial last workgroup) gives the same bits for them
    if rows > 130:
        assert net.set_fused(True)
        np.testing.assert_array_equal(net.forward(x[:rows - 129]), y_fused[:rows - 129])


def test_fused_not_used_for_other_shapes(gpu):
    ce = gpu
    assert not ce.MLP([7, 64, 50, 3]).set_fused(True)
    assert not ce.MLP([100, 256, 256, 8]).set_fused(True)  # input wider than 62 at H = 256
    assert not ce.MLP([63, 256, 256, 8]).set_fused(True)   # no room for the bias columns
    assert ce.MLP([62, 256, 256, 8]).set_fused(True)
    assert not ce.MLP([40] + [256] * 6 + [8]).set_fused(True)  # 7 weight layers
    assert not ce.MLP([40] + [512] * 5 + [8]).set_fused(True)  # 6 at H = 512
    assert not ce.MLP([130, 512, 512, 8]).set_fused(True)  # input wider than 128 at H = 512
