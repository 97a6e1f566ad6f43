"""The allocation-free gradient entry point (xk_mlp_grad_noalloc, partials in a
fixed device array of the library) against xk_mlp_grad with its workspace
argument, through the C-ABI: bitwise equal gradients and loss on two calls in a
row, the second with fewer rows (so fewer gradient workgroups, leaving stale
partial rows of the first call in the array), and an invalid call rejected."""

import ctypes

import numpy as np
import pytest

from conftest import pkg


def _data(torch, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, (n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.2).astype(np.float32)
    return torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()


@pytest.mark.gpu
def test_noalloc_gradient_equals_workspace_gradient(gpu):
    torch = gpu
    L = pkg("_lib")
    lib = L.lib()
    n_params = int(lib.xk_mlp_n_params())
    rng = np.random.default_rng(11)
    params = torch.from_numpy((0.2 * rng.normal(size=n_params)).astype(np.float32)).cuda()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n, seed in ((20000, 1), (777, 2)):
        print(f"no-alloc gradient: n={n} seed={seed}", flush=True)
        X, y = _data(torch, n, seed)
        work = torch.empty(int(lib.xk_mlp_grad_work(n)), dtype=torch.float32, device="cuda")
        ref = torch.empty(n_params + 1, dtype=torch.float32, device="cuda")
        got = torch.full((n_params + 1,), float("nan"), dtype=torch.float32, device="cuda")
        L.check(lib.xk_mlp_grad(L.ptr(X), L.ptr(y), n, L.ptr(params), 3.0, L.ptr(ref), L.ptr(work),
                                work.numel(), stream), "xk_mlp_grad")
        L.check(lib.xk_mlp_grad_noalloc(L.ptr(X), L.ptr(y), n, L.ptr(params), 3.0, L.ptr(got), stream),
                "xk_mlp_grad_noalloc")
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(g))
        assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), float(np.max(np.abs(r - g)))
    assert lib.xk_mlp_grad_noalloc(L.ptr(X), L.ptr(y), 0, L.ptr(params), 3.0, L.ptr(got), stream) < 0
    assert b"invalid" in lib.xk_last_error()
