"""The CPU MLP oracle pinned by central finite differences (FP64)."""
import numpy as np


def test_backward_matches_finite_differences():
    from mlp_oracle import init_params, loss_grad
    fin, width, hidden, fout = 5, 7, 3, 4
    p = init_params(fin, width, hidden, fout, seed=3)
    p += 0.05 * np.random.default_rng(1).standard_normal(p.shape)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((9, fin))
    t = rng.standard_normal((9, fout))
    _, g, _ = loss_grad(p, x, t, fin, width, hidden, fout)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (loss_grad(p + e, x, t, fin, width, hidden, fout)[0] -
                 loss_grad(p - e, x, t, fin, width, hidden, fout)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_bf16_rounding_is_round_to_nearest_even():
    from mlp_oracle import bf16_round
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert bf16_round(np.float32(1.0 + 2 ** -8)) == 1.0
    assert bf16_round(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    assert bf16_round(np.float32(-2.5)) == -2.5
