"""Learned candidate prescreen: the network and its training on the GPU.

No counterpart in the reference.  An 8-64-64-1 ReLU network scores every
candidate pair at the start of an increment from features of the
increment-start geometry (``DeviceContext.lp_features``); the fused assembly
then runs its pair kernel only over the pairs the network keeps plus every
contributing pair, and the verification pass (``DeviceContext.lp_verify``, the exact
narrowphase over the rest at the end of the Newton call) counts the pruned
pairs that ended within contact range.  A miss makes the driver re-run the
call on the full path, so every pair that can enter contact keeps the full
path's state.

Training data are logged from the method's own runs: features at increment
start, label 1 for a non-contributing pair whose exact gap at the end of the
Newton call is within contact range (the event the verification pass counts).  The
forward pass, the gradient of the weighted logistic loss and Adam all run in
the library's kernels (``xk_mlp_forward``, ``xk_mlp_grad``, ``xk_mlp_adam``);
the threshold is then set below the lowest positive logit of the training
set by a margin, so that recall is 1 there.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .context import require_gpu

N_IN, N_H = 8, 64
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "prescreen_weights.npz")


def n_params() -> int:
    return int(L.lib().xk_mlp_n_params())


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases except a negative output bias (the
    positive class is rare)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / N_IN), (N_IN, N_H))
    W2 = rng.normal(0.0, np.sqrt(2.0 / N_H), (N_H, N_H))
    w3 = rng.normal(0.0, np.sqrt(1.0 / N_H), N_H)
    p = np.concatenate([W1.ravel(), np.zeros(N_H), W2.ravel(), np.zeros(N_H), w3, [-2.0]])
    assert p.size == n_params()
    return p.astype(np.float32)


class PrescreenNet:
    """Parameters (FP32, on the device) plus the keep threshold on the logit."""

    def __init__(self, params=None, *, threshold: float = 0.0, seed: int = 0):
        self.torch = torch = require_gpu()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.size != n_params():
            raise ValueError(f"expected {n_params()} parameters")
        self.params = torch.as_tensor(p.copy(), device="cuda")
        self.threshold = float(threshold)
        self._m = torch.zeros_like(self.params)
        self._v = torch.zeros_like(self.params)
        self._t = 0

    # ------------------------------------------------------------ kernels
    def _X(self, X):
        t = self.torch.as_tensor(X, dtype=self.torch.float32, device="cuda").contiguous()
        if t.ndim != 2 or t.shape[1] != N_IN:
            raise ValueError("features must be (n, 8)")
        return t

    def logits(self, X):
        X = self._X(X)
        out = self.torch.empty(X.shape[0], dtype=self.torch.float32, device="cuda")
        L.check(L.lib().xk_mlp_forward(L.ptr(X), X.shape[0], L.ptr(self.params), L.ptr(out),
                                       self.torch.cuda.current_stream().cuda_stream), "xk_mlp_forward")
        return out

    def loss_and_grad(self, X, y, pos_weight: float):
        X = self._X(X)
        y = self.torch.as_tensor(y, dtype=self.torch.float32, device="cuda").contiguous()
        n = X.shape[0]
        work = self.torch.empty(int(L.lib().xk_mlp_grad_work(n)), dtype=self.torch.float32, device="cuda")
        gl = self.torch.empty(n_params() + 1, dtype=self.torch.float32, device="cuda")
        L.check(L.lib().xk_mlp_grad(L.ptr(X), L.ptr(y), n, L.ptr(self.params), float(pos_weight), L.ptr(gl),
                                    L.ptr(work), work.numel(), self.torch.cuda.current_stream().cuda_stream),
                "xk_mlp_grad")
        return float(gl[-1].item()), gl[:-1]

    def adam_step(self, grad, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self._t += 1
        L.check(L.lib().xk_mlp_adam(L.ptr(self.params), L.ptr(grad), L.ptr(self._m), L.ptr(self._v), self._t,
                                    float(lr), float(b1), float(b2), float(eps),
                                    self.torch.cuda.current_stream().cuda_stream), "xk_mlp_adam")

    # ------------------------------------------------------------ training
    def fit(self, X, y, *, epochs: int = 300, lr: float = 3e-3, pos_weight: float | None = None,
            batch: int = 0, seed: int = 0):
        """Adam on the weighted logistic loss; full batch unless ``batch``.
        Returns the loss history."""
        X = self._X(X)
        y = self.torch.as_tensor(y, dtype=self.torch.float32, device="cuda")
        n = X.shape[0]
        if pos_weight is None:
            npos = float(y.sum().item())
            pos_weight = max(1.0, (n - npos) / max(npos, 1.0))
        rng = np.random.default_rng(seed)
        hist = []
        for _ in range(epochs):
            if batch and batch < n:
                idx = self.torch.as_tensor(rng.permutation(n)[:batch], device="cuda")
                loss, g = self.loss_and_grad(X[idx], y[idx], pos_weight)
            else:
                loss, g = self.loss_and_grad(X, y, pos_weight)
            self.adam_step(g, lr)
            hist.append(loss)
        return hist

    def choose_threshold(self, X, y, margin: float = 2.0) -> float:
        """Keep threshold = lowest logit of a positive sample minus ``margin``."""
        z = self.logits(X).cpu().numpy()
        y = np.asarray(y)
        pos = z[y > 0.5]
        self.threshold = float(pos.min() - margin) if pos.size else float(np.quantile(z, 0.99))
        return self.threshold

    # ------------------------------------------------------------ storage
    def host_params(self) -> np.ndarray:
        return self.params.cpu().numpy().astype(np.float32)

    def save(self, path: str = WEIGHTS):
        np.savez(path, params=self.host_params(), threshold=np.float64(self.threshold))

    @classmethod
    def load(cls, path: str = WEIGHTS) -> "PrescreenNet":
        d = np.load(path)
        return cls(d["params"], threshold=float(d["threshold"]))


def labels_of(features: np.ndarray, gap_end: np.ndarray, g_on: float = 0.0):
    """Training rows of one Newton call: the non-contributing pairs (feature 7
    = 0), label 1 if their exact gap at the end of the call is <= g_on."""
    keep = features[:, 7] < 0.5
    return features[keep], (gap_end[keep] <= g_on).astype(np.float32)
