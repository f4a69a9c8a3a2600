"""CPU pins of the learned warm start's FP64 oracle (oracle/mlp_oracle.py):
  * BF16 rounding equals torch's float32 -> bfloat16 conversion (round to
    nearest, ties to even), an independent implementation;
  * the hand-written backward pass equals central finite differences of
    the FP64 forward pass;
  * with BF16 weights the gradient is the FP64 gradient at the rounded
    weights (the straight-through convention of the HIP kernels)."""
import numpy as np
import torch

from oracle import mlp_oracle as M


def params(nin, seed):
    rng = np.random.default_rng(seed)
    P = nin * 64 + 64 + 64 * 64 + 64 + 64 + 1
    return (rng.standard_normal(P) * 0.2).astype(np.float32)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10 ** rng.uniform(-5, 5, 10000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 65504.0],
                                 dtype=np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16_round(x), ref)


def test_backward_matches_finite_differences():
    nin, B = 10, 7
    rng = np.random.default_rng(2)
    p = params(nin, 3).astype(np.float64)
    x = rng.standard_normal((B, nin))
    t = rng.uniform(-1, 1, B)
    _, g = M.loss_and_grad(p, x, t, round_weights=False)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, x, t, round_weights=False)
        lm, _ = M.loss_and_grad(p - e, x, t, round_weights=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_rounded_gradient_is_gradient_at_rounded_weights():
    nin, B = 10, 16
    rng = np.random.default_rng(4)
    p = params(nin, 5)
    x = rng.standard_normal((B, nin))
    t = rng.uniform(-1, 1, B)
    W1, b1, W2, b2, W3, b3 = M.rounded(p, nin)
    pr = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])
    l1, g1 = M.loss_and_grad(p, x, t, round_weights=True)
    l2, g2 = M.loss_and_grad(pr, x, t, round_weights=False)
    assert l1 == l2 and np.array_equal(g1, g2)
