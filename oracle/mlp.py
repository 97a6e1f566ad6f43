"""Oracle: the candidate-prescreen MLP (8 -> 64 -> 64 -> 1, ReLU), FP64.

Test infrastructure only (see ``oracle/__init__.py``).  The network has no
counterpart in the reference; this module states what the device kernels
compute so that their forward pass and their gradients can be checked:

* parameters, flat FP32, in the order W1 (8x64, row k = input), b1 (64),
  W2 (64x64, row k = hidden-1 unit), b2 (64), w3 (64), b3 (1): 4801 values;
* BF16 is round-to-nearest-even of the FP32 bit pattern (the top 16 bits);
* training forward: inputs and W1, W2 rounded to BF16, everything else FP32
  values, accumulation exact here (FP64); loss = mean over samples of the
  weighted logistic loss  w+ y softplus(-z) + (1 - y) softplus(z);
  gradients flow straight through the BF16 rounding of the weights;
* inference forward: as training, and the first hidden activations rounded
  to BF16 before the second layer (what the MFMA kernel feeds its matrix
  cores);
* Adam (Kingma & Ba 2015) with bias correction.
"""

from __future__ import annotations

import numpy as np

N_IN, N_H = 8, 64
SIZES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("w3", (N_H,)), ("b3", (1,))]
N_PARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16_round(x) -> np.ndarray:
    """FP32 values rounded to the nearest BF16 (ties to even), returned as FP32."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32)


def softplus(z):
    return np.logaddexp(0.0, z)


def sigmoid(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def forward(params, X, *, inference: bool = False):
    """Logits z3 (n,) and the intermediates of the training forward."""
    P = unpack(np.asarray(params, np.float32))
    W1 = bf16_round(P["W1"]).astype(np.float64)
    W2 = bf16_round(P["W2"]).astype(np.float64)
    x = bf16_round(X).astype(np.float64)
    z1 = x @ W1 + P["b1"].astype(np.float64)
    a1 = np.maximum(z1, 0.0)
    a1_in = bf16_round(a1.astype(np.float32)).astype(np.float64) if inference else a1
    z2 = a1_in @ W2 + P["b2"].astype(np.float64)
    a2 = np.maximum(z2, 0.0)
    z3 = a2 @ P["w3"].astype(np.float64) + float(P["b3"][0])
    return z3, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, W2=W2, w3=P["w3"].astype(np.float64))


def loss_and_grad(params, X, y, pos_weight: float):
    """Mean weighted logistic loss and its gradient w.r.t. the FP32 parameters."""
    y = np.asarray(y, np.float64)
    n = len(y)
    z3, c = forward(params, X)
    loss = float(np.mean(pos_weight * y * softplus(-z3) + (1.0 - y) * softplus(z3)))
    g3 = (-pos_weight * y * sigmoid(-z3) + (1.0 - y) * sigmoid(z3)) / n
    d2 = (g3[:, None] * c["w3"][None, :]) * (c["z2"] > 0.0)
    d1 = (d2 @ c["W2"].T) * (c["z1"] > 0.0)
    g = dict(W1=c["x"].T @ d1, b1=d1.sum(0), W2=c["a1"].T @ d2, b2=d2.sum(0), w3=c["a2"].T @ g3,
             b3=np.array([g3.sum()]))
    return loss, pack(g)


def adam_step(params, grad, m, v, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (t counts from 1); returns new (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1**t)
    vh = v / (1.0 - b2**t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
