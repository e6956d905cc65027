"""BF16 MFMA kernels of the warm-start MLP against the FP64 CPU oracle.

The oracle emulates the device's BF16 rounding of weights, inputs and
stored activations, so forward outputs agree to FP32-vs-FP64 accumulation
(3e-3) and gradients to 3e-2 relative L2 (a BF16 rounding that falls the
other way at a tie region propagates through the backward pass).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("width,hidden,rows", [(64, 2, 128), (128, 4, 256), (256, 6, 2048)])
def test_forward_and_gradient_match_oracle(width, hidden, rows):
    from dgbgk.mlp import MLP
    from mlp_oracle import init_params, loss_grad
    print(f"mlp width={width} hidden={hidden} rows={rows}", flush=True)
    m = MLP(fin=5, width=width, hidden=hidden, fout=4, batch=max(rows, 64))
    p = init_params(5, width, hidden, 4, seed=width + hidden).astype(np.float32)
    p += (0.05 * np.random.default_rng(7).standard_normal(p.shape)).astype(np.float32)
    m.set_params(p)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((rows, 5)).astype(np.float32)
    t = rng.standard_normal((rows, 4)).astype(np.float32)
    loss, g = m.loss_grad(x, t)
    lo, go, yo = loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                           5, width, hidden, 4, bf16=True)
    y = m.forward(x)
    assert _rel(y, yo) < 3e-3
    assert abs(loss - lo) <= 3e-3 * lo
    assert _rel(g, go) < 3e-2


def test_training_reduces_loss():
    from dgbgk.mlp import MLP
    print("mlp training", flush=True)
    rng = np.random.default_rng(0)
    n = 20000
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] ** 2 - x[:, 3], np.exp(-x[:, 4] ** 2),
                  x[:, 0] * x[:, 4]], axis=1).astype(np.float32)
    m = MLP(fin=5, width=128, hidden=4, fout=4, batch=1024)
    m.init(seed=1)
    m.set_data(x, y)
    losses = m.train(600, lr=2e-3, seed=5)
    assert losses[-50:].mean() < 0.05 * losses[:10].mean()
    assert np.all(np.isfinite(losses))
