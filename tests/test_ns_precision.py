"""Why the sign iteration's products run on FP32 MFMA and not BF16 (CPU).

The projection W_+ = W (I + sign(W)) / 2 of the chosen form runs the 13
quintic steps of madam_ns_coef.h (three products each) on FP32 MFMA, then
one FP64 Newton step S (3 I - S^2) / 2 restores S^2 = I.  That finish fixes
the eigenvalues of S, not its eigenvectors: an error made in an early step
rotates the invariant subspaces and stays.  Emulated here with numpy (BF16 =
FP32 rounded to 8 mantissa bits, round to nearest even; sums in FP64, the
product rounded to FP32 as an MFMA accumulator would hold it):

  products in FP32                          ~4e-7 max|W| from eigh's W_+
  BF16 in steps 0-3, FP32 after             ~2e-2
  split BF16 (hi*hi + hi*lo + lo*hi) all    ~6e-5

BF16 throughput (2.5 PFLOP/s dense on MI355X, 16x FP32) would matter only
for batches (a single node's products are latency-bound), and neither form
keeps the bound within 1e-6 of the FP64 oracle.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
COEF_H = os.path.join(HERE, "..", "madam_amd", "csrc", "madam_ns_coef.h")


def coefficients():
    txt = open(COEF_H).read()
    body = txt[txt.index("MADAM_NS_COEF_INIT"):]
    vals = [float(x) for x in re.findall(r"(-?\d+\.\d+)f", body)]
    return np.array(vals).reshape(-1, 3)


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def product(A, B, mode):
    f64 = lambda X, Y: X.astype(np.float64) @ Y.astype(np.float64)
    if mode == "f32":
        return f64(A, B).astype(np.float32)
    if mode == "bf16":
        return f64(bf16(A), bf16(B)).astype(np.float32)
    Ah, Bh = bf16(A), bf16(B)
    Al, Bl = bf16(A - Ah), bf16(B - Bh)
    return (f64(Ah, Bh) + f64(Ah, Bl) + f64(Al, Bh)).astype(np.float32)


def projection(W, modes):
    n = W.shape[0]
    I = np.eye(n, dtype=np.float32)
    S = (W / np.linalg.norm(W)).astype(np.float32)
    for (a, b, c), m in zip(coefficients(), modes):
        T = product(S, S, m)
        U = product(T, (b * I + c * T).astype(np.float32), m)
        S = product(S, (a * I + U).astype(np.float32), m)
    S = S.astype(np.float64)
    S = S @ (3.0 * np.eye(n) - S @ S) / 2.0
    return W @ (np.eye(n) + S) / 2.0


def test_bf16_products_lose_the_projection():
    assert coefficients().shape == (13, 3)
    rng = np.random.default_rng(0)
    n = 100
    M = rng.standard_normal((n, n))
    W = (M + M.T) / 2
    w, V = np.linalg.eigh(W)
    exact = (V * np.maximum(w, 0.0)) @ V.T
    scale = np.abs(W).max()

    def err(modes):
        return np.abs(projection(W, modes) - exact).max() / scale

    e32 = err(["f32"] * 13)
    e_bf_early = err(["bf16"] * 4 + ["f32"] * 9)
    e_split = err(["split"] * 13)
    print(f"FP32 {e32:.1e}, BF16 in steps 0-3 {e_bf_early:.1e}, split BF16 {e_split:.1e}")
    assert e32 <= 1e-5
    assert e_bf_early >= 1e-3
    assert e_split >= 5 * e32
