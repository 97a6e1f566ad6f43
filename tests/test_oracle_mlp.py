"""The prescreen MLP oracle (oracle/mlp.py) pinned against independent
implementations: torch's BF16 conversion, torch autograd (float64, the BF16
weight rounding as a straight-through constant), central finite differences,
and torch.optim.Adam."""

import numpy as np
import pytest

from oracle import mlp

torch = pytest.importorskip("torch")


def test_bf16_rounding_matches_torch_and_closed_forms():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000)])
    x = x.astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(mlp.bf16_round(x), ref)
    # ties to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert mlp.bf16_round(np.float32(1 + 2**-8)) == 1.0
    assert mlp.bf16_round(np.float32(1 + 3 * 2**-8)) == np.float32(1 + 2**-6)
    assert mlp.bf16_round(np.float32(-2.0)) == -2.0


def torch_loss(p, X, y, pw):
    P = {}
    o = 0
    for name, shape in mlp.SIZES:
        n = int(np.prod(shape))
        P[name] = p[o:o + n].reshape(shape)
        o += n
    rb = lambda t: (t + (torch.tensor(mlp.bf16_round(t.detach().numpy().astype(np.float32)),
                                      dtype=torch.float64) - t).detach())
    x = torch.tensor(mlp.bf16_round(X), dtype=torch.float64)
    a1 = torch.relu(x @ rb(P["W1"]) + P["b1"])
    a2 = torch.relu(a1 @ rb(P["W2"]) + P["b2"])
    z = a2 @ P["w3"] + P["b3"][0]
    yt = torch.tensor(y, dtype=torch.float64)
    return torch.mean(pw * yt * torch.nn.functional.softplus(-z) + (1 - yt) * torch.nn.functional.softplus(z))


def case(seed=1, n=200):
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 0.5, mlp.N_IN * mlp.N_H), rng.normal(0, 0.1, mlp.N_H),
                        rng.normal(0, 0.2, mlp.N_H * mlp.N_H), rng.normal(0, 0.1, mlp.N_H),
                        rng.normal(0, 0.3, mlp.N_H), [0.1]]).astype(np.float32)
    X = rng.normal(0, 1, (n, mlp.N_IN)).astype(np.float32)
    y = (rng.random(n) < 0.2).astype(np.float64)
    return p, X, y


def test_gradient_matches_torch_autograd():
    p, X, y = case()
    loss, g = mlp.loss_and_grad(p, X, y, 7.0)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    lt = torch_loss(pt, X, y, 7.0)
    lt.backward()
    assert loss == pytest.approx(lt.item(), rel=1e-12)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_gradient_of_unrounded_parameters_matches_finite_differences():
    """b1, b2, w3, b3 enter unrounded: central differences pin their gradient."""
    p, X, y = case(2)
    _, g = mlp.loss_and_grad(p, X, y, 3.0)
    o = mlp.N_IN * mlp.N_H
    idx = [o + 3, o + 40, o + mlp.N_H + mlp.N_H * mlp.N_H + 5, mlp.N_PARAM - 10, mlp.N_PARAM - 1]
    for k in idx:
        h = 1e-3
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp.loss_and_grad(pp, X, y, 3.0)[0] - mlp.loss_and_grad(pm, X, y, 3.0)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-9)


def test_adam_matches_torch():
    p, X, y = case(3)
    _, g = mlp.loss_and_grad(p, X, y, 2.0)
    m = np.zeros(mlp.N_PARAM)
    v = np.zeros(mlp.N_PARAM)
    pn, m, v = mlp.adam_step(p.astype(np.float64), g, m, v, 1, 1e-2)
    pt = torch.nn.Parameter(torch.tensor(p.astype(np.float64)))
    opt = torch.optim.Adam([pt], lr=1e-2)
    pt.grad = torch.tensor(g)
    opt.step()
    np.testing.assert_allclose(pn, pt.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_inference_forward_rounds_hidden_activations():
    p, X, _ = case(4, 50)
    zt, c = mlp.forward(p, X)
    zi, _ = mlp.forward(p, X, inference=True)
    assert np.abs(zi - zt).max() <= 1e-2 * (1 + np.abs(zt).max())
    a1b = mlp.bf16_round(c["a1"].astype(np.float32)).astype(np.float64)
    P = mlp.unpack(p)
    z2 = a1b @ mlp.bf16_round(P["W2"]).astype(np.float64) + P["b2"]
    assert np.allclose(zi, np.maximum(z2, 0) @ P["w3"] + P["b3"][0], rtol=1e-12, atol=1e-12)
