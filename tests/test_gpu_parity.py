"""GPU parity of the gfx950 assembly path against the reference's own outputs
(golden fixtures) and the FP64 oracle, through the C-ABI library.

Bars: FP64 mode -- R and K within 1e-10 relative (Frobenius), identical
candidate and active sets, identical pair statuses; mixed mode -- residual
within 1e-6 relative, identical active set."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import golden, pkg
from oracle import assembly as oasm
from oracle import beam as obeam
from oracle import contact as ocon

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_R_TOL = 1e-6
MIXED_K_TOL = 1e-5


def csr(g, p):
    return sp.csr_matrix((g[p + "_data"], g[p + "_indices"], g[p + "_indptr"]), shape=tuple(g[p + "_shape"]))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def srel(A, B):
    return float(sp.linalg.norm(A - B) / sp.linalg.norm(B))


def section(r):
    C = pkg("context")
    return C.circle_section(r, 200e3, 0.3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_beam_force_and_tangent_match_reference(gpu, precision):
    g = golden("elem_small.npz")
    E = pkg("elements")
    s = section(float(g["radius"]))
    asm = E.ULCRBeamAssembler(g["X"], g["conn"], s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y, s.kappa_z,
                              precision=precision)
    for k in range(3):
        f = asm.assemble_internal_force(g[f"u{k}"])
        K = asm.assemble_tangent(g[f"u{k}"])
        assert rel(f, g[f"f{k}"]) < 1e-12, k
        assert srel(K, csr(g, f"K{k}")) < (FP64_TOL if precision == "fp64" else MIXED_K_TOL), k


def test_beam_matches_oracle_at_37_wires(gpu):
    wl = pkg("workloads")
    E = pkg("elements")
    st = wl.make_state(wl.C2)
    s = section(wl.C2.wire_radius)
    asm = E.ULCRBeamAssembler(st.mesh.node_coords, st.mesh.connectivity, s.E, s.G, s.A, s.Iy, s.Iz, s.J,
                              s.kappa_y, s.kappa_z)
    osec = obeam.circle_section(wl.C2.wire_radius, 200e3, 0.3)
    Kr, fr = obeam.assemble_cr_beam3d(st.mesh.node_coords, st.mesh.connectivity, st.u, osec)
    assert rel(asm.assemble_internal_force(st.u), fr) < FP64_TOL
    assert srel(asm.assemble_tangent(st.u), Kr) < FP64_TOL


def test_broadphase_is_the_exact_overlap_set(gpu):
    wl = pkg("workloads")
    C = pkg("context")
    st = wl.make_state(wl.C2)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, section(m.wire_radius), radii=m.radii)
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    ref = ocon.broadphase_aabb(x[m.connectivity[:, 0]], x[m.connectivity[:, 1]], m.radii)
    for cell in (None, 0.2, 1.0, 3.0):
        got = ctx.broadphase(ctx.to_dev(x), cell_size=cell)
        assert np.array_equal(got, ref), cell


def test_detect_candidates_matches_reference_sets(gpu):
    a = golden("assembly_7wire.npz")
    C = pkg("context")
    K = pkg("contact")
    r = float(a["radius"])
    radii = np.full(len(a["conn"]), r)
    ctx = C.DeviceContext(a["X"], a["conn"], section(r), radii=radii)
    mgr = K.ContactManager(ctx, K.ContactConfig(mu=0.3))
    out = K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, a["X"], a["conn"], radii))
    assert np.array_equal(np.array(out.candidates).reshape(-1, 2), a["cand0"])
    x = a["X"] + a["inc0_it0_u"].reshape(-1, 6)[:, :3]
    out = K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, x, a["conn"], radii))
    assert np.array_equal(np.array(out.candidates).reshape(-1, 2), a["inc0_cand"])


def compare_pairs(d, a, pre, tol):
    key = np.stack([np.minimum(d["elem_a"], d["elem_b"]), np.maximum(d["elem_a"], d["elem_b"])], 1)
    o = np.lexsort((key[:, 1], key[:, 0]))
    assert np.array_equal(key[o], a[pre + "_key"])
    assert np.array_equal(d["status"][o], a[pre + "_status"]), pre
    assert np.array_equal(d["stick"][o].astype(bool), a[pre + "_stick"]), pre
    for f in ["s", "t", "gap", "normal", "tangent1", "tangent2", "p_n", "z_t", "k_t", "q_trial_norm"]:
        r = a[pre + "_" + f]
        err = np.abs(d[f][o] - r).max() / max(np.abs(r).max(), 1e-300)
        assert err < tol, (pre, f, err)


@pytest.mark.parametrize("precision,staged", [("fp64", False), ("mixed", False), ("fp64", True), ("mixed", True)])
def test_fused_newton_assembly_matches_reference_two_increments(gpu, precision, staged):
    """Two increments x two Newton attempts of a 7-wire strand with Huber contact
    and Coulomb friction, against the reference's ContactForceAssembly /
    TangentAssembly / GeneralizedAlpha outputs, on both element paths."""
    a = golden("assembly_7wire.npz")
    C, K, T, A = pkg("context"), pkg("contact"), pkg("time_integration"), pkg("assembly")
    k_pen, sd, mu, rho_inf, dt, rho, ktr = a["params"]
    r = float(a["radius"])
    X, conn = a["X"], a["conn"]
    radii = np.full(len(conn), r)
    ctx = C.DeviceContext(X, conn, section(r), radii=radii, precision=precision, staged=staged)
    assert ctx.direct == (not staged)
    mgr = K.ContactManager(ctx, K.ContactConfig(mu=mu, k_t_ratio=ktr))
    K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, X, conn, radii))
    E = pkg("elements")
    s = section(r)
    Md = E.ULCRBeamAssembler(X, conn, s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y, s.kappa_z).lumped_mass_diag(rho)
    ga = T.GeneralizedAlphaProcess(Md, rho_inf=rho_inf)
    asm = A.StrandAssembly(ctx, k_pen=k_pen, smoothing_delta=sd, k_t_ratio=ktr, mu=mu, f_ext=a["f_ext"],
                           fixed_dofs=a["fixed"], config=mgr.config, time=ga, dt=dt)
    torch = ctx.torch
    u = torch.zeros(ctx.ndof, dtype=torch.float64, device=ctx.dev)
    state_tol = 1e-12 if precision == "fp64" else 1e-9
    for inc in range(2):
        u_pred = ga.predict_device(u, dt).cpu().numpy()
        assert rel(u_pred, a[f"inc{inc}_u_pred"]) < 1e-14
        asm.begin_increment(a[f"inc{inc}_it0_u"])
        compare_pairs(ctx.get_pairs(), a, f"inc{inc}_start", state_tol)
        for it in range(2):
            R, Kv, res2, nact = asm.assemble(a[f"inc{inc}_it{it}_u"])
            pre = f"inc{inc}_it{it}"
            Rh = R.cpu().numpy()
            assert rel(Rh, a[pre + "_R"]) < (FP64_TOL if precision == "fp64" else MIXED_R_TOL), pre
            assert float(res2.item()) == pytest.approx(float(Rh @ Rh), rel=1e-12)
            assert srel(asm.K_csr(), csr(a, pre + "_K")) < (FP64_TOL if precision == "fp64" else MIXED_K_TOL), pre
            assert int(nact.item()) == int(a[pre + "_n_active"])
            compare_pairs(ctx.get_pairs(), a, pre + "_pairs", state_tol)
        u = asm.u.clone()
        asm.set_u_ref(u)


def test_assembly_leaves_the_attempts_ga_state_unless_told_not_to(gpu):
    """By default the fused assembly leaves the attempt's acceleration and
    velocity in the time integrator, as the reference's loop does with
    GeneralizedAlphaProcess.correct before effective_residual
    (time_integration/strategy.py:178-188, _newton_dynamic.py:190-191):
    acc = c0 (u - u_pred), vel = v_pred + dt gamma acc.  With ga_state=False
    (a driver that corrects at the increment's end) it leaves them alone, and
    R and K are bitwise the same."""
    a = golden("assembly_7wire.npz")
    C, K, T, A, E = pkg("context"), pkg("contact"), pkg("time_integration"), pkg("assembly"), pkg("elements")
    k_pen, sd, mu, rho_inf, dt, rho, ktr = a["params"]
    r = float(a["radius"])
    X, conn = a["X"], a["conn"]
    radii = np.full(len(conn), r)
    s = section(r)
    Md = E.ULCRBeamAssembler(X, conn, s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y, s.kappa_z).lumped_mass_diag(rho)
    out = []
    for ga_state in (True, False):
        ctx = C.DeviceContext(X, conn, section(r), radii=radii, precision="fp64")
        mgr = K.ContactManager(ctx, K.ContactConfig(mu=mu, k_t_ratio=ktr))
        K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, X, conn, radii))
        ga = T.GeneralizedAlphaProcess(Md, rho_inf=rho_inf)
        asm = A.StrandAssembly(ctx, k_pen=k_pen, smoothing_delta=sd, k_t_ratio=ktr, mu=mu, f_ext=a["f_ext"],
                               fixed_dofs=a["fixed"], config=mgr.config, time=ga, dt=dt, ga_state=ga_state)
        torch = ctx.torch
        ga.predict_device(torch.zeros(ctx.ndof, dtype=torch.float64, device=ctx.dev), dt)
        asm.begin_increment(a["inc0_it0_u"])
        ga.d_acc.fill_(7.0)
        ga.d_vel.fill_(-7.0)
        R, Kv, _, _ = asm.assemble(a["inc0_it1_u"])
        u, up, vp = (x.cpu().numpy() for x in (asm.u, ga.d_u_pred, ga.d_v_pred))
        acc, vel = ga.d_acc.cpu().numpy(), ga.d_vel.cpu().numpy()
        if ga_state:
            c0 = 1.0 / (ga.beta * dt**2)
            acc_ref = c0 * (u - up)
            np.testing.assert_allclose(acc, acc_ref, rtol=1e-14, atol=0)
            np.testing.assert_allclose(vel, vp + dt * ga.gamma * acc_ref, rtol=1e-14, atol=1e-300)
            assert np.abs(acc_ref).max() > 0
        else:
            assert np.all(acc == 7.0) and np.all(vel == -7.0)
        out.append((R.cpu().numpy(), Kv.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("staged", [False, True])
def test_strategy_api_matches_oracle_on_37_wires(gpu, staged):
    """Drop-in Processes / Strategies (DetectCandidates, UpdateGeometry, Huber,
    Coulomb evaluate + tangent) against the oracle, one call each."""
    strategy_roundtrip("C2", staged)


@pytest.mark.parametrize("sd,mu", [(0.0, 0.3), (0.0, 0.0), (None, 0.0)],
                         ids=["plain_penalty", "plain_penalty_frictionless", "huber_frictionless"])
def test_strategy_api_plain_penalty_and_frictionless(gpu, sd, mu):
    """The same calls with smoothing_delta = 0 (p_n = max(0, k_pen (-gap)),
    contact_force/strategy.py:126-170) and with mu = 0 (no friction force,
    no friction tangent), on the 7-wire strand."""
    strategy_roundtrip("C1", False, smoothing_delta=sd, mu=mu)


def strategy_roundtrip(case_name, staged, smoothing_delta=None, mu=None):
    wl = pkg("workloads")
    C, K = pkg("context"), pkg("contact")
    case = getattr(wl, case_name)
    sd = case.smoothing_delta if smoothing_delta is None else smoothing_delta
    mu = case.mu if mu is None else mu
    st = wl.make_state(case)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, section(m.wire_radius), radii=m.radii, staged=staged)
    E = pkg("elements")
    s = section(m.wire_radius)
    beam = E.ULCRBeamAssembler(m.node_coords, m.connectivity, s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y,
                               s.kappa_z, context=ctx)
    cfg = K.ContactConfig(mu=mu)
    mgr = K.ContactManager(ctx, cfg)
    ocfg = ocon.ContactConfig(mu=mu)
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, ocfg)
    K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, m.node_coords, m.connectivity, m.radii))
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    op, ocand = ocon.detect_candidates(op, x, m.connectivity, m.radii, ocfg)
    out = K.DetectCandidatesProcess().process(K.DetectCandidatesInput(mgr, x, m.connectivity, m.radii))
    assert np.array_equal(np.array(out.candidates).reshape(-1, 2), ocand)
    op = ocon.update_geometry(op, x, ocfg)
    K.UpdateGeometryProcess().process(K.UpdateGeometryInput(mgr, x))
    d = ctx.get_pairs()
    assert np.array_equal(d["elem_a"], op.elem_a) and np.array_equal(d["elem_b"], op.elem_b)
    assert np.array_equal(d["status"], op.status)
    ndof = ctx.ndof
    hub = K.HuberContactForceProcess(ndof, smoothing_delta=sd)
    fc, res = hub.evaluate(st.u, mgr, case.k_pen)
    ofc, ores, op = ocon.huber_evaluate(op, case.k_pen, sd, ndof)
    assert rel(fc, ofc) < FP64_TOL and rel(res, ores) < FP64_TOL
    fr = K.CoulombReturnMappingProcess(ndof, k_pen=case.k_pen, k_t_ratio=0.5)
    ff, fres = fr.evaluate(st.u, mgr.pairs, mu, u_ref=st.u_ref)
    ofr = ocon.friction_evaluate(op, st.u, st.u_ref, case.k_pen, 0.5, mu, ndof)
    assert rel(ff, ofr.f_fric) < FP64_TOL
    # residual entries max(0, |q| - mu p_n) are rounding-level for slipping pairs:
    # compare on the scale of the friction bound mu p_n
    assert fres.shape == ofr.residual.shape
    if fres.size:
        assert np.abs(fres - ofr.residual).max() <= 1e-12 * mu * ofr.pairs.p_n.max()
    assert len(fr.friction_tangents) == int(ofr.mask.sum())
    Kc = hub.tangent(st.u, mgr, case.k_pen)
    assert srel(Kc, ocon.huber_tangent(ofr.pairs, case.k_pen, sd, ndof)) < FP64_TOL
    Kf = fr.tangent(st.u, mgr.pairs, mu)
    oKf = ocon.friction_tangent(ofr.pairs, ofr, ndof)
    if mu > 0.0:
        assert srel(Kf, oKf) < FP64_TOL
    else:
        assert abs(Kf).sum() == 0.0 and abs(oKf).sum() == 0.0 and not np.any(ff)
    d = ctx.get_pairs()
    assert np.array_equal(d["status"], ofr.pairs.status)
    # beam callbacks sharing the context (pattern now holds contact blocks too)
    osec = obeam.circle_section(case.wire_radius, 200e3, 0.3)
    Kr, fr_ = obeam.assemble_cr_beam3d(m.node_coords, m.connectivity, st.u, osec)
    assert rel(beam.assemble_internal_force(st.u), fr_) < FP64_TOL
    assert srel(beam.assemble_tangent(st.u), Kr) < FP64_TOL


def _rot(v, axis, ang):
    """Rodrigues rotation of v about the unit axis."""
    axis = axis / np.linalg.norm(axis)
    return v * np.cos(ang) + np.cross(axis, v) * np.sin(ang) + axis * (axis @ v) * (1 - np.cos(ang))


def test_narrowphase_and_frame_branches_match_oracle(gpu):
    """Hand-built segment pairs through every branch of the narrowphase
    (_compute.py:55-151): crossing, parallel overlapping and disjoint, s
    clamped, t clamped then s recomputed, endpoint to endpoint, collinear,
    zero-length A and B, intersecting (zero distance: fallback normal),
    near-parallel.  Each geometry meets five previous frames (_compute.py:154-261):
    none, a rotated previous normal (transport), the same normal (no
    rotation), a previous tangent without a normal, and a previous tangent
    along the new normal (the axis fallback).  Statuses alternate, so the
    hysteresis switches both ways.  FP64 s, t, gap, normal, tangent1 and
    status against the oracle's update_geometry."""
    C = pkg("context")
    L = pkg("_lib")
    geoms = [
        ((-1, 0, 0), (1, 0, 0), (0, -1, 0.3), (0, 1, 0.3)),          # crossing
        ((0, 0, 0), (1, 0, 0), (0.2, 0.3, 0), (1.2, 0.3, 0)),         # parallel, overlapping
        ((0, 0, 0), (1, 0, 0), (2, 0.3, 0), (3, 0.3, 0)),             # parallel, disjoint
        ((0, 0, 0), (1, 0, 0), (-0.5, -1, 0.2), (-0.4, 1, 0.25)),     # s clamped
        ((0, -1, 0), (0.1, 1, 0), (-2, 0.2, 0.3), (-1, 0.25, 0.3)),   # t clamped, s recomputed
        ((0, 0, 0), (1, 0, 0), (1.5, 0.5, 0.1), (2.5, 1.5, 0.1)),     # endpoint to endpoint
        ((0, 0, 0), (1, 0, 0), (0.5, 0, 0), (1.5, 0, 0)),             # collinear (parallel branch: s = 0)
        ((0, 0, 0), (1, 0, 0), (0.3, 0.2, 0.1), (0.3, 0.2, 0.1)),     # zero-length B
        ((0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0, 0, 0), (1, 0, 0)),     # zero-length A
        ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)),               # intersecting
        ((0, 0, 0), (1, 0, 0), (0, 0.3, 0), (1, 0.3, 1e-4)),          # near-parallel
    ]
    X = np.array([pt for g in geoms for pt in g], float)
    conn = np.arange(len(X)).reshape(-1, 2)
    r = np.full(len(conn), 0.1)
    ocfg = ocon.ContactConfig()
    ea0 = 2 * np.arange(len(geoms))
    eb0 = ea0 + 1
    # the new normal of each geometry (fresh pairs), to build the previous frames from
    n_new = ocon.update_geometry(ocon.Pairs.fresh(ea0, eb0, conn, r, r), X, ocfg).normal
    ea, eb, tp, npv = [], [], [], []
    for k in range(len(geoms)):
        n = n_new[k]
        perp = np.cross(n, [0.3, 0.5, 0.8])
        perp /= np.linalg.norm(perp)
        n_rot = _rot(n, perp + 0.5 * np.cross(n, perp), 0.35)
        t_rot = np.cross(n_rot, [0.6, -0.2, 0.7])
        t_rot /= np.linalg.norm(t_rot)
        for t1, nn in ((np.zeros(3), np.zeros(3)), (t_rot, n_rot), (perp, n), (perp, np.zeros(3)), (n, n)):
            ea.append(ea0[k]); eb.append(eb0[k]); tp.append(t1); npv.append(nn)
    ea, eb, tp, npv = np.array(ea), np.array(eb), np.array(tp), np.array(npv)
    status = (np.arange(len(ea)) % 2).astype(np.int8)
    op = ocon.Pairs.fresh(ea, eb, conn, r, r)
    op.tangent1, op.normal, op.status = tp.copy(), npv.copy(), status.astype(np.int64)
    op = ocon.update_geometry(op, X, ocfg)
    ctx = C.DeviceContext(X, conn, section(0.1), radii=r)
    ctx.set_pairs(ea, eb, state={"tangent1": tp, "normal": npv, "status": status})
    ctx.pairs_run(L.PAIR_GEOM | L.PAIR_HYST, g_on=ocfg.g_on, g_off=ocfg.g_off)
    d = ctx.get_pairs()
    for name in ("s", "t", "gap", "normal", "tangent1"):
        np.testing.assert_allclose(d[name], getattr(op, name), rtol=0, atol=1e-12, err_msg=name)
    assert np.array_equal(d["status"], op.status)
    # the branches were met: zero-distance fallback normal, clamps at both ends, both statuses
    assert np.all(d["normal"][9 * 5:10 * 5] == [0.0, 0.0, 1.0])
    assert {0.0, 1.0} <= set(np.round(d["s"], 12)) and {0.0, 1.0} <= set(np.round(d["t"], 12))
    assert set(d["status"].tolist()) == {0, 1}
    # UpdateGeometryInput(allow_deactivation=False) and (freeze_active_set=True)
    for kw, modes in (({"allow_deactivation": False}, L.PAIR_GEOM | L.PAIR_HYST), ({"freeze_active_set": True}, L.PAIR_GEOM)):
        ref = ocon.Pairs.fresh(ea, eb, conn, r, r)
        ref.tangent1, ref.normal, ref.status = tp.copy(), npv.copy(), status.astype(np.int64)
        ref = ocon.update_geometry(ref, X, ocfg, **kw)
        ctx.set_pairs(ea, eb, state={"tangent1": tp, "normal": npv, "status": status})
        ctx.pairs_run(modes, g_on=ocfg.g_on, g_off=ocfg.g_off, allow_deactivation=kw.get("allow_deactivation", True))
        d = ctx.get_pairs()
        assert np.array_equal(d["status"], ref.status), kw
        np.testing.assert_allclose(d["tangent1"], ref.tangent1, rtol=0, atol=1e-12)
    assert np.array_equal(d["status"], status)      # frozen: statuses as set


def test_non_chain_connectivity_takes_the_staged_path(gpu):
    """Reversed and reordered elements: the library detects that the mesh is
    not made of element chains and assembles through the staged path."""
    wl = pkg("workloads")
    C, A = pkg("context"), pkg("assembly")
    case = wl.StrandCase("shuffled_7wire", 7, 1.0, 20.0, 20, 10.0, seed=11)
    st = wl.make_state(case)
    m = st.mesh
    conn = m.connectivity.copy()
    conn[20:40] = conn[20:40, ::-1]  # wire 1 runs backwards
    perm = np.random.default_rng(3).permutation(len(conn))
    conn = conn[perm]
    ctx = C.DeviceContext(m.node_coords, conn, section(case.wire_radius), radii=m.radii)
    assert not ctx.direct
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=st.f_ext, fixed_dofs=st.fixed_dofs)
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.begin_increment(st.u)
    R, Kv, _, nact = asm.assemble(st.u)
    osec = obeam.circle_section(case.wire_radius, 200e3, 0.3)
    cfg = ocon.ContactConfig(mu=case.mu)
    prob = oasm.Problem(m.node_coords, conn, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, conn, m.radii, cfg)
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    op, _ = ocon.detect_candidates(op, x, conn, m.radii, cfg)
    op = ocon.update_geometry(op, x, cfg)
    ref = oasm.assemble_iteration(prob, op, st.u, st.u_ref, st.f_ext, st.fixed_dofs, None, 0.0)
    assert rel(R.cpu().numpy(), ref.R) < FP64_TOL
    assert srel(asm.K_csr(), ref.K) < FP64_TOL
    assert int(nact.item()) == ref.n_active


@pytest.mark.parametrize("layers", [True, False], ids=["between_layers", "all_pairs"])
def test_mixed_precision_residual_and_active_set_at_91_wires(gpu, layers):
    """Full-size 91-wire workload: FP64 kernels against the oracle (1e-10) and
    the mixed form against FP64 (residual 1e-6, identical active set).  With
    the same-layer filter (the benchmark's candidate set) and without it (every
    non-adjacent pair, including overlapping elements of one wire: many more
    active pairs)."""
    wl = pkg("workloads")
    C, A, T, E, Kc = pkg("context"), pkg("assembly"), pkg("time_integration"), pkg("elements"), pkg("contact")
    case = wl.C3
    st = wl.make_state(case)
    m = st.mesh
    lmap = wl.layer_map(st.elem_layer) if layers else None
    s = section(case.wire_radius)
    Md = E.ULCRBeamAssembler(m.node_coords, m.connectivity, s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y,
                             s.kappa_z).lumped_mass_diag(case.rho)
    dt = case.dt()
    outs = {}
    for prec in ("fp64", "mixed", "fp32"):
        ctx = C.DeviceContext(m.node_coords, m.connectivity, s, radii=m.radii, precision=prec)
        ga = T.GeneralizedAlphaProcess(Md, rho_inf=case.rho_inf)
        ga.predict_device(ctx.torch.zeros(ctx.ndof, dtype=ctx.torch.float64, device=ctx.dev), dt)
        asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                               f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, time=ga, dt=dt,
                               config=Kc.ContactConfig(mu=case.mu, exclude_same_layer=layers, elem_layer_map=lmap))
        asm.begin_increment(np.zeros(ctx.ndof))  # candidates on the reference configuration
        asm.begin_increment(st.u)
        R, Kv, res2, nact = asm.assemble(st.u)
        outs[prec] = (R.cpu().numpy(), asm.K_csr(), int(nact.item()), ctx.get_pairs()["status"].copy())
    # oracle
    osec = obeam.circle_section(case.wire_radius, 200e3, 0.3)
    cfg = ocon.ContactConfig(mu=case.mu, exclude_same_layer=layers, elem_layer_map=lmap)
    prob = oasm.Problem(m.node_coords, m.connectivity, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, cfg)
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    op, _ = ocon.detect_candidates(op, x, m.connectivity, m.radii, cfg)
    op = ocon.update_geometry(op, x, cfg)
    ga = oasm.GeneralizedAlpha(obeam.lumped_mass_diag(m.node_coords, m.connectivity, case.rho, osec), case.rho_inf)
    ga.predict(np.zeros_like(st.u), dt)
    ref = oasm.assemble_iteration(prob, op, st.u, st.u_ref, st.f_ext, st.fixed_dofs, ga, dt)
    R64, K64, n64, s64 = outs["fp64"]
    assert rel(R64, ref.R) < FP64_TOL
    assert srel(K64, ref.K) < FP64_TOL
    assert n64 == ref.n_active
    assert np.array_equal(s64, ref.pairs.status)
    for prec in ("mixed", "fp32"):
        Rm, Km, nm, sm = outs[prec]
        assert rel(Rm, R64) < MIXED_R_TOL, prec
        assert srel(Km, K64) < MIXED_K_TOL, prec
        assert nm == n64 and np.array_equal(sm, s64), prec


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_instances_match_oracle_and_never_pair_across(gpu, precision):
    """Three 7-wire instances of different seeds concatenated into one context."""
    wl = pkg("workloads")
    C, A = pkg("context"), pkg("assembly")
    case = wl.C1
    b = wl.concat_states([wl.make_state(case, seed=s) for s in (7, 8, 9)])
    ctx = C.DeviceContext(b.node_coords, b.connectivity, section(case.wire_radius), radii=b.radii,
                          precision=precision)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=b.f_ext, fixed_dofs=b.fixed_dofs)
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.begin_increment(b.u)
    R, Kv, res2, nact = asm.assemble(b.u)
    d = ctx.get_pairs()
    inst = np.searchsorted([e1 for _, e1 in b.elem_ranges], d["elem_a"], side="right")
    inst_b = np.searchsorted([e1 for _, e1 in b.elem_ranges], d["elem_b"], side="right")
    assert np.array_equal(inst, inst_b) and len(set(inst.tolist())) == 3
    osec = obeam.circle_section(case.wire_radius, 200e3, 0.3)
    cfg = ocon.ContactConfig(mu=case.mu)
    prob = oasm.Problem(b.node_coords, b.connectivity, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    op, _ = ocon.detect_candidates(ocon.Pairs(), b.node_coords, b.connectivity, b.radii, cfg)
    x = b.node_coords + b.u.reshape(-1, 6)[:, :3]
    op, _ = ocon.detect_candidates(op, x, b.connectivity, b.radii, cfg)
    op = ocon.update_geometry(op, x, cfg)
    ref = oasm.assemble_iteration(prob, op, b.u, b.u_ref, b.f_ext, b.fixed_dofs, None, 0.0)
    Rh = R.cpu().numpy()
    for n0, n1 in b.node_ranges:
        sl = slice(6 * n0, 6 * n1)
        assert rel(Rh[sl], ref.R[sl]) < (FP64_TOL if precision == "fp64" else MIXED_R_TOL)
    assert srel(asm.K_csr(), ref.K) < (FP64_TOL if precision == "fp64" else MIXED_K_TOL)
    assert int(nact.item()) == ref.n_active
    assert np.array_equal(ctx.get_pairs()["status"], ref.pairs.status)


def test_assembly_is_bitwise_deterministic_and_graph_replay_matches(gpu):
    wl = pkg("workloads")
    C, A = pkg("context"), pkg("assembly")
    case = wl.C1
    st = wl.make_state(case)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, section(case.wire_radius), radii=m.radii)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=st.f_ext, fixed_dofs=st.fixed_dofs)
    asm.begin_increment(st.u)
    snap = ctx.get_pairs()
    R1, K1, _, n1 = (t.clone() for t in asm.assemble(st.u))
    # restore the pair state the first call consumed, then repeat
    ctx.set_pairs(snap["elem_a"], snap["elem_b"], snap)
    asm.K = ctx.new_K()
    R2, K2, _, n2 = (t.clone() for t in asm.assemble(st.u))
    assert ctx.torch.equal(R1, R2) and ctx.torch.equal(K1, K2) and int(n1.item()) == int(n2.item())
    # graph: replay several times with eager work between, each against an eager run
    asm.capture()
    for rep in range(3):
        ctx.set_pairs(snap["elem_a"], snap["elem_b"], snap)
        ctx.pattern()
        Rg, Kg, _, ng = (t.clone() for t in asm.replay())
        ctx.torch.cuda.synchronize()
        assert ctx.torch.equal(Rg, R1) and ctx.torch.equal(Kg, K1), rep
        _ = (ctx.X * 2.0).sum().item()
