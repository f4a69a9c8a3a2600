"""Learned warm start for the node ADMM (no reference counterpart).

A node's cost matrix L (PP->L) -> pair features (madam_pair_features, GPU)
-> the MLP (madam_mlp_forward, GPU) predicts the converged primal entries
X*_ij -> madam_warm_start turns the prediction into (X, y, Q) for
ADMM_solver.  Training targets are the primal matrices the FP64 ADMM
(plain form: Jacobi projection, exact cut solve) converges to from a cold
start on the training graphs; training runs on the GPU (madam_mlp_train).
The bound never depends on the start: ADMM_solver's safe bound is
certified for whatever (y, s) it ends at, and the cold start stays
available as the fallback.
"""
from __future__ import annotations

import numpy as np

import madam_amd as G


def cold_start(L):
    n = L.shape[0]
    return np.asfortranarray(np.eye(n)), np.zeros(n), np.zeros((n, n), order="F")


def solve_basic(Ls, tol, starts=None, projection=0, itermax=20000, batch=64):
    """Basic SDP (no cuts) by the GPU ADMM for each L, in batches.
    Returns the AdmmResults (X, y, f, nbit, ...)."""
    out = []
    for b0 in range(0, len(Ls), batch):
        probs = []
        for q in range(b0, min(len(Ls), b0 + batch)):
            L = Ls[q]
            X, y, Q = starts[q] if starts is not None else cold_start(L)
            probs.append(dict(L=L, X=np.array(X, order="F"), y=np.array(y), Q=np.array(Q, order="F"),
                              sigma=1.0, tol=tol, itermax=itermax))
        out.extend(G.ADMM_solver_batch(probs, projection=projection))
    return out


def pair_values(X):
    n = X.shape[0]
    iu = np.triu_indices(n, 1)
    return X[iu].astype(np.float32)


class LearnedWarmStart:
    def __init__(self, params, mu, sd):
        self.params = np.asarray(params, np.float32)
        self.mu = np.asarray(mu, np.float32)
        self.sd = np.asarray(sd, np.float32)

    @classmethod
    def train(cls, Ls, Xs, epochs=2, batch=4096, lr=2e-3, seed=0):
        """Fit on graphs Ls with converged primals Xs; returns (model, losses)."""
        raw = np.concatenate([G.pair_features(L) for L in Ls])
        mu = raw.mean(axis=0)
        sd = raw.std(axis=0)
        sd[sd == 0] = 1.0
        feat = ((raw - mu) / sd).astype(np.float32)
        tgt = np.concatenate([pair_values(X) for X in Xs])
        params, losses = G.mlp_train(G.mlp_init(G.NFEAT, seed), feat, tgt, epochs=epochs,
                                     batch=batch, lr=lr, seed=seed)
        return cls(params, mu, sd), losses

    def predict(self, L):
        return G.mlp_forward(self.params, G.pair_features(L, self.mu, self.sd))

    def start(self, L):
        return G.warm_start(L, self.predict(L))

    def save(self, path):
        np.savez(path, params=self.params, mu=self.mu, sd=self.sd)

    @classmethod
    def load(cls, path):
        d = np.load(path)
        return cls(d["params"], d["mu"], d["sd"])
