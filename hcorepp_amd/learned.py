"""Learned rank predictor of the batched recompression.

A small MLP (features -> 64 -> 64 -> 1, ReLU) predicts log2 of the post-truncation rank of an
update from features the batched kernel computes on the device right after stacking (input
ranks, tile size, accuracy, and log-norms of the leading diagonal entries of the R factors of the
two blocks of the stacked V factor).  It is trained on the GPU with the library's own FP32
forward / backward / Adam kernels (``hcorepp_mlp_*``) on recompressions the method itself
performed (features from ``Batch.features()``, labels = the exact ranks), and runs inside the
batched kernel with BF16 weights (``Batch.set_rank_model``), where it presizes the core handed
to the Jacobi SVD; an exact test on the pivoted R factor accepts the presized core or falls back
to the exact one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import DeviceBuffer, _check, lib


def n_params() -> int:
    return int(lib().hcorepp_rank_model_params())


def n_features() -> int:
    return int(lib().hcorepp_rank_model_features())


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised FP32 parameters in the library's layout W1, b1, W2, b2, w3, b3."""
    rng = np.random.default_rng(seed)
    f, h = n_features(), 64
    w1 = rng.standard_normal((f, h)) * np.sqrt(2.0 / f)
    w2 = rng.standard_normal((h, h)) * np.sqrt(2.0 / h)
    w3 = rng.standard_normal(h) * np.sqrt(1.0 / h)
    p = np.concatenate([w1.ravel(), np.zeros(h), w2.ravel(), np.zeros(h), w3, [5.0]]).astype(np.float32)
    assert p.size == n_params()
    return p


class RankPredictor:
    """Device-resident FP32 parameters plus Adam state; trains with the library's kernels."""

    def __init__(self, ctx, params: np.ndarray | None = None, seed: int = 0):
        self.ctx = ctx
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.params = DeviceBuffer.from_numpy(p)
        self.m = DeviceBuffer.from_numpy(np.zeros_like(p))
        self.v = DeviceBuffer.from_numpy(np.zeros_like(p))
        self.grads = DeviceBuffer(p.nbytes)
        self.step = 0
        self._bf16 = None

    def host_params(self) -> np.ndarray:
        return self.params.download(np.float32, n_params())

    def grad(self, X: np.ndarray, y: np.ndarray):
        """(loss, gradient) of the mean-squared error on (X, y), computed on the device."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = X.shape[0]
        dX, dy = DeviceBuffer.from_numpy(X), DeviceBuffer.from_numpy(y)
        ws = DeviceBuffer(4 * lib().hcorepp_mlp_workspace_floats(n))
        loss = C.c_float(0)
        _check(lib().hcorepp_mlp_grad_f32(self.ctx.h, self.params.ptr, dX.ptr, dy.ptr, n, ws.ptr, self.grads.ptr,
                                          C.byref(loss)))
        return float(loss.value), self.grads.download(np.float32, n_params())

    def train(self, X: np.ndarray, y: np.ndarray, epochs: int = 200, batch: int = 4096, lr: float = 2e-3,
              seed: int = 0, log=None):
        """Adam on minibatches of (X, y) (device-resident), all arithmetic in the library's kernels."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = X.shape[0]
        rng = np.random.default_rng(seed)
        ws = DeviceBuffer(4 * lib().hcorepp_mlp_workspace_floats(batch))
        dX = DeviceBuffer(4 * batch * X.shape[1])
        dy = DeviceBuffer(4 * batch)
        L = lib()
        hist = []
        for ep in range(epochs):
            perm = rng.permutation(n)
            tot, cnt = 0.0, 0
            for s0 in range(0, n - batch + 1, batch):
                idx = perm[s0:s0 + batch]
                dX.upload(X[idx])
                dy.upload(y[idx])
                loss = C.c_float(0)
                _check(L.hcorepp_mlp_grad_f32(self.ctx.h, self.params.ptr, dX.ptr, dy.ptr, batch, ws.ptr,
                                              self.grads.ptr, C.byref(loss)))
                self.step += 1
                _check(L.hcorepp_mlp_adam_f32(self.ctx.h, self.params.ptr, self.grads.ptr, self.m.ptr, self.v.ptr,
                                              lr, 0.9, 0.999, 1e-8, self.step))
                tot += loss.value
                cnt += 1
            hist.append(tot / max(cnt, 1))
            if log and (ep % 20 == 0 or ep == epochs - 1):
                log(f"[rank-mlp] epoch {ep}: train mse {hist[-1]:.4f}")
        self._bf16 = None
        return hist

    def bf16(self) -> DeviceBuffer:
        """BF16 copy of the parameters for the in-kernel inference."""
        if self._bf16 is None:
            self._bf16 = DeviceBuffer(2 * n_params())
            _check(lib().hcorepp_mlp_pack_bf16(self.ctx.h, self.params.ptr, self._bf16.ptr))
            self.ctx.sync()
        return self._bf16

    def predict_log2(self, X: np.ndarray) -> np.ndarray:
        """The in-kernel BF16 inference on rows of X (log2 rank)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        dX = DeviceBuffer.from_numpy(X)
        out = DeviceBuffer(4 * max(n, 1))
        _check(lib().hcorepp_mlp_predict_bf16(self.ctx.h, self.bf16().ptr, dX.ptr, n, out.ptr))
        self.ctx.sync()
        return out.download(np.float32, n)

    def save(self, path: str, **meta):
        np.savez(path, params=self.host_params(), **meta)

    @staticmethod
    def load(ctx, path: str) -> "RankPredictor":
        with np.load(path, allow_pickle=False) as z:
            return RankPredictor(ctx, params=z["params"])
