"""CPU FP64 restatement of the learned warm start's network (test
infrastructure only: the values madam_amd/csrc/madam_mlp.hip is compared with).

Same architecture and conventions as the HIP kernels: weights rounded to
BF16 (nearest, ties to even) before use, biases unrounded, tanh at every
layer, loss mean((o - t)^2 / 2), gradients with respect to the rounded
weights.  Pinned by finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np

H = 64


def bf16_round(x):
    """Nearest BF16, ties to even, as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32)
    return np.where(special, a, out)


def unpack(params, nin):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + nin * H].reshape(nin, H); o += nin * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def rounded(params, nin):
    """Parameters with the weights rounded to BF16 (biases kept)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, nin)
    r = lambda w: bf16_round(w).astype(np.float64)
    return r(W1), b1, r(W2), b2, r(W3), b3


def forward(params, x, nin=None, round_weights=True):
    nin = nin or x.shape[1]
    W1, b1, W2, b2, W3, b3 = rounded(params, nin) if round_weights else unpack(params, nin)
    x = np.asarray(x, np.float64)
    h1 = np.tanh(x @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    o = np.tanh(h2 @ W3 + b3)
    return o, (x, h1, h2)


def loss_and_grad(params, x, t, round_weights=True):
    nin = x.shape[1]
    W1, b1, W2, b2, W3, b3 = rounded(params, nin) if round_weights else unpack(params, nin)
    o, (x, h1, h2) = forward(params, x, nin, round_weights)
    t = np.asarray(t, np.float64)
    B = x.shape[0]
    loss = np.mean(0.5 * (o - t) ** 2)
    dz3 = (o - t) / B * (1 - o ** 2)
    gW3 = h2.T @ dz3
    gb3 = dz3.sum()
    da2 = np.outer(dz3, W3) * (1 - h2 ** 2)
    gW2 = h1.T @ da2
    gb2 = da2.sum(axis=0)
    da1 = (da2 @ W2.T) * (1 - h1 ** 2)
    gW1 = x.T @ da1
    gb1 = da1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
