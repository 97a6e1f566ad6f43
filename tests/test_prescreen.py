"""Learned candidate prescreen: the network kernels against the oracle
(oracle/mlp.py), exactness of the prescreened assembly, and what the trained
network gains over the method without it."""

import numpy as np
import pytest

from conftest import pkg
from oracle import mlp

pytestmark = pytest.mark.gpu


def rand_case(seed=1, n=3000):
    rng = np.random.default_rng(seed)
    p = mlp.pack(dict(W1=rng.normal(0, 0.6, (8, 64)), b1=rng.normal(0, 0.2, 64), W2=rng.normal(0, 0.2, (64, 64)),
                      b2=rng.normal(0, 0.1, 64), w3=rng.normal(0, 0.3, 64), b3=[0.2])).astype(np.float32)
    X = rng.normal(0, 1, (n, 8)).astype(np.float32)
    y = (rng.random(n) < 0.1).astype(np.float32)
    return p, X, y


def test_mfma_forward_matches_oracle(gpu):
    P = pkg("prescreen")
    for seed, n in ((1, 3000), (2, 33), (3, 1)):
        p, X, _ = rand_case(seed, n)
        net = P.PrescreenNet(p)
        z = net.logits(X).cpu().numpy()
        zo, _ = mlp.forward(p, X, inference=True)
        assert z.shape == zo.shape
        assert np.abs(z - zo).max() <= 1e-2 * max(1.0, np.abs(zo).max()), (seed, np.abs(z - zo).max())


def test_gradient_matches_oracle(gpu):
    P = pkg("prescreen")
    for seed, n in ((4, 5000), (5, 70)):
        p, X, y = rand_case(seed, n)
        net = P.PrescreenNet(p)
        loss, g = net.loss_and_grad(X, y, 6.0)
        lo, go = mlp.loss_and_grad(p, X, y, 6.0)
        assert loss == pytest.approx(lo, rel=1e-5)
        g = g.cpu().numpy()
        assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
        # bitwise reproducible
        assert np.array_equal(net.loss_and_grad(X, y, 6.0)[1].cpu().numpy(), g)


def test_adam_matches_oracle(gpu):
    P = pkg("prescreen")
    p, X, y = rand_case(6, 500)
    net = P.PrescreenNet(p)
    _, g = net.loss_and_grad(X, y, 3.0)
    net.adam_step(g, 1e-2)
    pn, _, _ = mlp.adam_step(p.astype(np.float64), g.cpu().numpy().astype(np.float64), np.zeros(p.size),
                             np.zeros(p.size), 1, 1e-2)
    np.testing.assert_allclose(net.host_params(), pn, rtol=1e-5, atol=1e-7)


def test_training_separates_a_learnable_rule(gpu):
    """Label = first feature below a cut: the network trained by the library's
    kernels reaches high recall and precision."""
    P = pkg("prescreen")
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 3] < -1.2).astype(np.float32)
    net = P.PrescreenNet(seed=3)
    hist = net.fit(X, y, epochs=400, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0]
    z = net.logits(X).cpu().numpy()
    pred = z >= 0.0
    recall = (pred & (y > 0.5)).sum() / (y > 0.5).sum()
    precision = (pred & (y > 0.5)).sum() / max(pred.sum(), 1)
    assert recall > 0.97 and precision > 0.7


def run(case, precision="fp64", net=None, log=False):
    return pkg("train_prescreen").run_case(case, precision=precision, net=net, log=log)


def training_set(runs):
    P = pkg("prescreen")
    Xs, ys = [], []
    for _, b in runs:
        for feat, gap in b.feature_log:
            x, y = P.labels_of(feat, gap)
            Xs.append(x)
            ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


PAIR_FIELDS = ("elem_a", "elem_b", "s", "t", "gap", "normal", "tangent1", "p_n", "k_pen", "k_t", "q_trial_norm",
               "dissipation", "z_t", "status", "stick")


def same_pair_state(ctx_a, ctx_b, mask=None):
    """Every state field of every pair bitwise equal (frames and friction history included)."""
    a, b = ctx_a.get_pairs(), ctx_b.get_pairs()
    for k in PAIR_FIELDS:
        x, y = a[k], b[k]
        if mask is not None:
            x, y = x[mask], y[mask]
        assert np.array_equal(x, y), k


def same_trajectory(a, b):
    assert len(a.calls) == len(b.calls)
    for p, q in zip(a.calls, b.calls):
        assert (p.load_frac, p.n_attempts, p.n_active, p.converged) == (q.load_frac, q.n_attempts, q.n_active,
                                                                         q.converged)
        assert p.res_history == q.res_history     # bitwise: the contributing pairs are always evaluated
    assert np.array_equal(a.u, b.u)


def test_prescreen_keeps_the_trajectory_exact_and_prunes(gpu):
    """Train on features logged from full-path runs of three frictionless
    squeezes, then solve held-out squeezes with the prescreen, frictionless and
    with Coulomb friction mu = 0.3: identical trajectory bit for bit (attempts,
    active counts, residual histories, final u), every pair's final state bit
    for bit (the verification pass replays, for the pruned pairs, the frame
    transport of every assembly of the call), and the share of pairs the
    narrowphase visits per Newton attempt against the method without it (all
    candidates)."""
    P = pkg("prescreen")
    W = pkg("workloads")
    train = [run(W.squeeze_case(f), log=True) for f in (80.0, 100.0, 120.0)]
    X, y = training_set(train)
    assert y.sum() > 0
    net = P.PrescreenNet(seed=0)
    hist = net.fit(X, y, epochs=300, lr=3e-3)
    assert hist[-1] < hist[0]
    net.choose_threshold(X, y, margin=1.0)
    cases = (W.squeeze_case(110.0), W.squeeze_case(100.0, n_wires=19),
             W.squeeze_case(60.0, mu=0.3), W.squeeze_case(100.0, n_wires=19, mu=0.3))
    for case in cases:
        full, bf = run(case)
        lp, bl = run(case, net=net)
        same_trajectory(full, lp)
        same_pair_state(bf.ctx, bl.ctx)   # frames of pruned pairs transported once per assembly, as the full path
        watched = sum(w for w, _ in lp.watch) / sum(n for _, n in lp.watch)
        print(f"watched {watched:.3f} of candidate pairs per attempt, rechecks {lp.n_rechecks} "
              f"of {len(lp.calls)} calls")
        assert watched < 0.5
        assert lp.n_rechecks <= len(lp.calls) // 4


def test_prescreened_assembly_is_bitwise_the_full_path_at_91_wires(gpu):
    """With the committed weights on the 91-wire workload: R, K_T and the active
    count of one assembly over the compacted watch list equal the full path's
    bit for bit, and so does every pair's state after the verification pass
    (watched pairs written back, pruned pairs' geometry replayed)."""
    wl, C, A, Kc, P = pkg("workloads"), pkg("context"), pkg("assembly"), pkg("contact"), pkg("prescreen")
    case = wl.C3
    st = wl.make_state(case)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, C.circle_section(case.wire_radius, case.E, case.nu),
                          radii=m.radii, precision="mixed")
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu, f_ext=st.f_ext,
                           fixed_dofs=st.fixed_dofs,
                           config=Kc.ContactConfig(mu=case.mu, exclude_same_layer=True,
                                                   elem_layer_map=wl.layer_map(st.elem_layer)))
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.set_prescreen(P.PrescreenNet.load(), elem_layer=st.elem_layer)
    asm.begin_increment(st.u)
    n_watch = asm.n_watch
    assert 0 < n_watch < ctx.n_pairs
    ctx.pairs_snapshot()
    R1, K1, _, n1 = (t.clone() for t in asm.assemble())
    assert ctx.lp_verify(asm.u) >= 0      # watched pairs written back, pruned pairs' geometry replayed
    lp_state = ctx.get_pairs()
    ctx.pairs_restore()
    R0, K0, _, n0 = (t.clone() for t in asm.assemble())     # full path from the same state
    full_state = ctx.get_pairs()
    assert gpu.equal(R0, R1) and gpu.equal(K0, K1) and int(n0.item()) == int(n1.item())
    for k in PAIR_FIELDS:
        assert np.array_equal(lp_state[k], full_state[k]), k
    print(f"91 wires: watched {n_watch} of {ctx.n_pairs} pairs")


def test_prescreen_at_1000_wires_is_the_full_path(gpu):
    """The 1000-wire strand with the committed weights, FP32-mixed: three
    assemblies at different displacements over the watch list (R, K_T and the
    active count of each bitwise the full path's), then the verification pass;
    every pair's state equals the full path's after the same three assemblies."""
    wl, C, A, Kc, P = pkg("workloads"), pkg("context"), pkg("assembly"), pkg("contact"), pkg("prescreen")
    case = wl.C5
    st = wl.make_state(case)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, C.circle_section(case.wire_radius, case.E, case.nu),
                          radii=m.radii, precision="mixed")
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu, f_ext=st.f_ext,
                           fixed_dofs=st.fixed_dofs,
                           config=Kc.ContactConfig(mu=case.mu, exclude_same_layer=True,
                                                   elem_layer_map=wl.layer_map(st.elem_layer)))
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.set_prescreen(P.PrescreenNet.load(), elem_layer=st.elem_layer)
    asm.begin_increment(st.u)
    n_watch = asm.n_watch
    print(f"1000 wires: watched {n_watch} of {ctx.n_pairs} pairs", flush=True)
    assert 0 < n_watch < ctx.n_pairs
    rng = np.random.default_rng(1001)
    us = [st.u + rng.normal(0.0, 2e-4 * case.wire_radius, st.u.shape) for _ in range(3)]
    ctx.pairs_snapshot()
    lp_out = [tuple(t.clone() for t in asm.assemble(u)) for u in us]
    assert ctx.lp_verify(asm.u) >= 0
    lp_state = ctx.get_pairs()
    ctx.pairs_restore()
    full_out = [tuple(t.clone() for t in asm.assemble(u)) for u in us]
    full_state = ctx.get_pairs()
    for (R1, K1, _, n1), (R0, K0, _, n0) in zip(lp_out, full_out):
        assert gpu.equal(R0, R1) and gpu.equal(K0, K1) and int(n0.item()) == int(n1.item())
    for k in PAIR_FIELDS:
        assert np.array_equal(lp_state[k], full_state[k]), k


def test_prescreen_under_device_newton_control(gpu):
    """The committed network on two frictional squeezes (mu = 0.3) solved with
    the Newton attempts and the driver's tests on the device (attempts enqueued
    back to back, the attempt after a call's end a no-op): the trajectory, the
    final u and every pair's final state equal the full path's under the same
    device control bit for bit, so the no-op attempts record no assembly for
    the verification pass's replay."""
    P, W, TP = pkg("prescreen"), pkg("workloads"), pkg("train_prescreen")
    net = P.PrescreenNet.load()
    for case in (W.squeeze_case(60.0, mu=0.3), W.squeeze_case(100.0, n_wires=19, mu=0.3)):
        print(f"prescreen under device control, {len(case['conn'])} elements", flush=True)
        full, bf = TP.run_case(case, device_solve=True)
        lp, bl = TP.run_case(case, net=net, device_solve=True)
        same_trajectory(full, lp)
        same_pair_state(bf.ctx, bl.ctx)
        assert sum(w for w, _ in lp.watch) < sum(n for _, n in lp.watch)
