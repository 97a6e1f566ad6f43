"""Kelvin-Voigt coating terms on the device against the FP64 oracle
(oracle/coating.py, pinned by tests/test_oracle_coating.py), through the C-ABI.

Workload: the 7-wire strand (C1) and the 37-wire strand (C2) with their seeded
perturbation, coating stiffness 0.1 k_pen, damping 0.5 k_coat coat_dt,
coating mu 0.2, with the core radii equal to the contact radii (what the
reference's solver passes) and 0.9 of them (a coating layer: the Huber
penalty acts on the core gap, the coating on the outer overlap).  Two
increments: at 0.5 u (then the compression is committed as the previous
step's) and at u with u_ref = 0.5 u, so the dashpot and the coating friction
history are both exercised.

Bars: FP64 -- R and K within 1e-10 relative, identical statuses, coating
state (compression, history, stick) within 1e-10; mixed -- R within 1e-6, K
within 1e-5 of the oracle, identical active set."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import pkg
from oracle import assembly as oasm
from oracle import beam as obeam
from oracle import coating as ocoat
from oracle import contact as ocon

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def srel(A, B):
    return float(sp.linalg.norm(A - B) / sp.linalg.norm(B))


def coating_params(case):
    k = 0.1 * case.k_pen
    coat_dt = 0.05
    return dict(coating_stiffness=k, coating_damping=0.5 * k * coat_dt, coating_mu=0.2, coating_k_t_ratio=0.5), \
        coat_dt


def device_run(case, st, prec, core, kw, coat_dt, steps, **ctx_opts):
    """Device: increments at the displacement list `steps`; the last one's
    assembly (R, K csr, n_active, pairs, coating state, rebuilds)."""
    C, A, Kc = pkg("context"), pkg("assembly"), pkg("contact")
    m = st.mesh
    s = C.circle_section(case.wire_radius, case.E, case.nu)
    ctx = C.DeviceContext(m.node_coords, m.connectivity, s, radii=m.radii, precision=prec, **ctx_opts)
    cfg = Kc.ContactConfig(mu=case.mu, **kw)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, config=cfg, core_radii=core)
    asm.set_coat_dt(coat_dt)
    asm.begin_increment(np.zeros(ctx.ndof))      # candidates on the reference configuration
    u_ref = np.zeros(ctx.ndof)
    out = None
    for k, (u_start, u_eval) in enumerate(steps):
        asm.set_u_ref(u_ref)
        asm.begin_increment(u_start)
        R, K, res2, nact = asm.assemble(u_eval)
        out = (R.cpu().numpy(), asm.K_csr(), int(nact.item()), ctx.get_pairs(), ctx.get_coating())
        if k + 1 < len(steps):
            asm.commit_coating()
            u_ref = u_eval
    assert ctx.status_flags() & 8 == 0, "a coating contribution had no slot"
    return out + (asm.n_coat_rebuilds,)


def oracle_run(case, st, core, kw, coat_dt, steps):
    m = st.mesh
    osec = obeam.circle_section(case.wire_radius, case.E, case.nu)
    cfg = ocon.ContactConfig(mu=case.mu, **kw)
    prob = oasm.Problem(m.node_coords, m.connectivity, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    cr = m.radii if core is None else np.broadcast_to(core, m.radii.shape)
    # as the device's begin_increment at rest: candidates and geometry (hysteresis) on X
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, cfg, core_radii=cr)
    op = ocon.update_geometry(op, m.node_coords, cfg)
    u_ref = np.zeros(6 * len(m.node_coords))
    ref = None
    for k, (u_start, u_eval) in enumerate(steps):
        x = m.node_coords + u_start.reshape(-1, 6)[:, :3]
        op, _ = ocon.detect_candidates(op, x, m.connectivity, m.radii, cfg, core_radii=cr)
        op = ocon.update_geometry(op, x, cfg)
        ref = oasm.assemble_iteration(prob, op, u_eval, u_ref, st.f_ext, st.fixed_dofs, None, 0.0, coat_dt=coat_dt)
        op = ref.pairs
        if k + 1 < len(steps):
            op = ocoat.commit(op)
            u_ref = u_eval
    return ref


def order(d):
    key = np.stack([np.minimum(d["elem_a"], d["elem_b"]), np.maximum(d["elem_a"], d["elem_b"])], 1)
    return np.lexsort((key[:, 1], key[:, 0])), key


def compare(case, st, prec, core_frac, steps, **ctx_opts):
    kw, coat_dt = coating_params(case)
    core = None if core_frac is None else core_frac * st.mesh.radii
    R, K, nact, pairs, coat, rebuilds = device_run(case, st, prec, core, kw, coat_dt, steps, **ctx_opts)
    ref = oracle_run(case, st, core, kw, coat_dt, steps)
    cc = ref.pairs.coating_compression
    print(f"pairs {len(cc)} compressed {int((cc > 0).sum())} n_active {ref.n_active} rebuilds {rebuilds}")
    assert (cc > 0).sum() > 0
    o, key = order(pairs)
    okey = np.stack([np.minimum(ref.pairs.elem_a, ref.pairs.elem_b), np.maximum(ref.pairs.elem_a, ref.pairs.elem_b)], 1)
    oo = np.lexsort((okey[:, 1], okey[:, 0]))
    assert np.array_equal(key[o], okey[oo])
    assert nact == ref.n_active
    assert np.array_equal(pairs["status"][o], ref.pairs.status[oo])
    if prec == "fp64":
        assert rel(R, ref.R) < FP64_TOL
        assert srel(K, ref.K) < FP64_TOL
        p = ref.pairs
        for name, r in (("coating_compression", p.coating_compression), ("coating_compression_prev",
                        p.coating_compression_prev), ("coating_z_t", p.coating_z_t),
                        ("coating_q_trial_norm", p.coating_q_trial_norm)):
            d, rr = coat[name][o], r[oo]
            err = np.abs(d - rr).max() / max(np.abs(rr).max(), 1e-300)
            if not err < FP64_TOL:
                bad = np.nonzero(np.abs(d - rr).reshape(len(d), -1).max(axis=1) > FP64_TOL * np.abs(rr).max())[0]
                for q in bad[:8]:
                    print(name, q, "dev", d[q], "ref", rr[q], "cc", p.coating_compression[oo][q], "ccp",
                          p.coating_compression_prev[oo][q], "status", p.status[oo][q], "p_n", p.p_n[oo][q],
                          "stick dev/ref", coat["coating_stick"][o][q], p.coating_stick[oo][q])
                print("n bad", len(bad), "of", len(d))
            assert err < FP64_TOL, (name, err)
        assert np.array_equal(coat["coating_stick"][o].astype(bool), p.coating_stick[oo])
        np.testing.assert_allclose(coat["core_radius_a"][o], p.core_radius_a[oo], rtol=0, atol=0)
    else:
        assert rel(R, ref.R) < 1e-6
        assert srel(K, ref.K) < 1e-5
    return rebuilds


@pytest.mark.parametrize("core_frac", [None, 0.9], ids=["core_eq_radius", "coating_layer"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coating_assembly_matches_oracle_7wire(gpu, precision, core_frac):
    wl = pkg("workloads")
    st = wl.make_state(wl.C1)
    half = 0.5 * st.u
    compare(wl.C1, st, precision, core_frac, [(half, half), (st.u, st.u)])


@pytest.mark.parametrize("opts", [{"staged": True}, {"host_pattern": True}], ids=["staged_path", "host_pattern"])
def test_coating_on_the_other_assembly_paths(gpu, opts):
    """The staged element path (any connectivity) and the host-built pattern
    carry the coating terms as well (two increments, coating layer, FP64)."""
    wl = pkg("workloads")
    st = wl.make_state(wl.C1)
    half = 0.5 * st.u
    compare(wl.C1, st, "fp64", 0.9, [(half, half), (st.u, st.u)], **opts)
    zero = np.zeros_like(st.u)
    assert compare(wl.C1, st, "fp64", None, [(zero, st.u)], **opts) >= 1   # probe + rebuild on this path


def test_coating_assembly_matches_oracle_37wire(gpu):
    wl = pkg("workloads")
    st = wl.make_state(wl.C2)
    half = 0.5 * st.u
    compare(wl.C2, st, "fp64", 0.9, [(half, half), (st.u, st.u)])


@pytest.mark.parametrize("core_frac", [None, 0.9], ids=["core_eq_radius", "coating_layer"])
def test_coating_strategy_api_matches_oracle(gpu, core_frac):
    """The drop-in Processes and the coating Strategy, one call each:
    DetectCandidates with core radii, UpdateGeometry with the coating branch,
    then KelvinVoigtCoatingProcess.forces / stiffness / friction_forces /
    friction_stiffness against the oracle's; a second step after the
    compression is committed (dashpot rate and coating history carried)."""
    wl = pkg("workloads")
    C, K = pkg("context"), pkg("contact")
    case = wl.C1
    st = wl.make_state(case)
    m = st.mesh
    ndof = 6 * len(m.node_coords)
    kw, coat_dt = coating_params(case)
    core = None if core_frac is None else core_frac * m.radii
    ctx = C.DeviceContext(m.node_coords, m.connectivity, C.circle_section(case.wire_radius, case.E, case.nu),
                          radii=m.radii)
    cfg = K.ContactConfig(mu=case.mu, **kw)
    mgr = K.ContactManager(ctx, cfg)
    ocfg = ocon.ContactConfig(mu=case.mu, **kw)
    det, geo = K.DetectCandidatesProcess(), K.UpdateGeometryProcess()
    det.process(K.DetectCandidatesInput(mgr, m.node_coords, m.connectivity, m.radii, core_radii=core))
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, ocfg, core_radii=core)
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    det.process(K.DetectCandidatesInput(mgr, x, m.connectivity, m.radii, core_radii=core))
    op, _ = ocon.detect_candidates(op, x, m.connectivity, m.radii, ocfg, core_radii=core)
    strat = K._create_coating_strategy(coating_stiffness=cfg.coating_stiffness)
    assert isinstance(strat, K.KelvinVoigtCoatingProcess)
    assert isinstance(K._create_coating_strategy(coating_stiffness=0.0), K.NoCoatingProcess)
    u_ref = np.zeros(ndof)
    for step, (u, dt) in enumerate(((st.u, coat_dt), (1.05 * st.u, 0.5 * coat_dt))):
        x = m.node_coords + u.reshape(-1, 6)[:, :3]
        geo.process(K.UpdateGeometryInput(mgr, x))
        op = ocon.update_geometry(op, x, ocfg)
        o, key = order(ctx.get_pairs())
        okey = np.stack([np.minimum(op.elem_a, op.elem_b), np.maximum(op.elem_a, op.elem_b)], 1)
        oo = np.lexsort((okey[:, 1], okey[:, 0]))
        assert np.array_equal(key[o], okey[oo])
        assert np.array_equal(ctx.get_pairs()["status"][o], op.status[oo])
        coat = ctx.get_coating()
        cc = op.coating_compression
        assert (cc > 0).sum() > 0
        assert np.abs(coat["coating_compression"][o] - cc[oo]).max() <= FP64_TOL * cc.max()
        f = strat.forces(mgr.pairs, x, cfg, dt)
        assert rel(f, ocoat.forces(op, ocfg, dt, ndof)) < FP64_TOL
        mk, bk = ocoat.stiffness_blocks(op, ocfg, dt)
        assert srel(strat.stiffness(mgr.pairs, x, cfg, ndof, dt), ocon.blocks_to_csr(op, mk, bk, ndof)) < FP64_TOL
        ff = strat.friction_forces(mgr.pairs, x, cfg, u, u_ref)
        off, op = ocoat.friction_forces(op, ocfg, u, u_ref, ndof)
        assert rel(ff, off) < FP64_TOL
        mf, bf = ocoat.friction_stiffness_blocks(op, ocfg)
        assert srel(strat.friction_stiffness(mgr.pairs, x, cfg, ndof), ocon.blocks_to_csr(op, mf, bf, ndof)) < FP64_TOL
        coat = ctx.get_coating()
        zr = op.coating_z_t[oo]
        assert np.abs(coat["coating_z_t"][o] - zr).max() <= FP64_TOL * np.abs(zr).max()
        assert np.array_equal(coat["coating_stick"][o].astype(bool), op.coating_stick[oo])
        print(f"step {step}: pairs {len(cc)} compressed {int((cc > 0).sum())} slipping "
              f"{int((~op.coating_stick & (cc > 0)).sum())}")
        ctx.coating_commit()
        op = ocoat.commit(op)
        u_ref = u
    # coating mu = 0: no coating friction (strategy.py:217-219, :309-311), the normal terms unchanged
    cfg0 = K.ContactConfig(mu=case.mu, **dict(kw, coating_mu=0.0))
    assert not np.any(strat.friction_forces(mgr.pairs, x, cfg0, u, u_ref))
    assert strat.friction_stiffness(mgr.pairs, x, cfg0, ndof).nnz == 0
    ocfg0 = ocon.ContactConfig(mu=case.mu, **dict(kw, coating_mu=0.0))
    assert rel(strat.forces(mgr.pairs, x, cfg0, dt), ocoat.forces(op, ocfg0, dt, ndof)) < FP64_TOL


def test_pair_pressing_within_the_increment_gets_a_slot(gpu):
    """Active set frozen at the rest configuration (no pair in contact, no
    compression), assembly at the perturbed u where pairs overlap: the
    reference applies the coating force to those INACTIVE pairs, so the probe
    must find them without a slot and the pattern be rebuilt before the
    assembly."""
    wl = pkg("workloads")
    st = wl.make_state(wl.C1)
    zero = np.zeros_like(st.u)
    rebuilds = compare(wl.C1, st, "fp64", None, [(zero, st.u)])
    assert rebuilds >= 1


def squeeze_solver(coat_k_frac, mu_coat):
    """The 7-wire dynamic radial squeeze (the Newton-trajectory fixture's case)
    with the coating on: core radii 0.95 r, k_coat = coat_k_frac k_pen."""
    wl, S = pkg("workloads"), pkg("solver")
    c = wl.squeeze_case(100.0)
    kw = dict(coating_stiffness=coat_k_frac * c["k_pen"], coating_damping=0.0, coating_mu=mu_coat,
              coating_k_t_ratio=0.5)
    cfg = S.SolverConfig(**c["cfg"])
    return c, kw, cfg, 0.95 * c["radii"]


_ORACLE_SQUEEZE = {}


def coated_squeeze_oracle():
    """The coated squeeze's calls on the oracle assembly (computed once)."""
    if "res" not in _ORACLE_SQUEEZE:
        from _oracle_backend import OracleBackend

        S, C, E = pkg("solver"), pkg("context"), pkg("elements")
        c, kw, cfg, core = squeeze_solver(0.001, 0.2)
        X, conn, radii = c["X"], c["conn"], c["radii"]
        lmap = {int(e): int(k) for e, k in enumerate(c["layer"])}
        sec = C.circle_section(1.0, 200e3, 0.3)
        Md = E.ULCRBeamAssembler(X, conn, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y,
                                 sec.kappa_z).lumped_mass_diag(7.85e-9)
        osec = obeam.circle_section(1.0, 200e3, 0.3)
        ocfg = ocon.ContactConfig(mu=c["mu"], exclude_same_layer=True, elem_layer_map=lmap, **kw)
        ob = OracleBackend(X, conn, radii, osec, M_diag=Md, rho_inf=c["rho_inf"], k_pen=c["k_pen"],
                           smoothing_delta=c["sd"], mu=c["mu"], k_t_ratio=c["ktr"], cfg=ocfg)
        ob.fixed = c["fixed"]
        # the oracle's pairs take the core radii at creation (DetectCandidatesInput.core_radii)
        ob.detect = lambda u, _ob=ob: setattr(_ob, "pairs", ocon.detect_candidates(
            _ob.pairs, _ob._x(u), conn, radii, ocfg, core_radii=core)[0])
        _ORACLE_SQUEEZE["res"] = S.ContactFrictionSolver(ob, fixed_dofs=c["fixed"], f_ext_total=c["f"],
                                                         t_total=c["t_total"], config=cfg).solve()
    return _ORACLE_SQUEEZE["res"]


def coated_squeeze_device(mode):
    """The coated squeeze on the device in `mode` -> (result, #compressed pairs at
    each commit, pattern rebuilds during each Newton call)."""
    S, C, Kc, E = pkg("solver"), pkg("context"), pkg("contact"), pkg("elements")
    c, kw, cfg, core = squeeze_solver(0.001, 0.2)
    X, conn, radii = c["X"], c["conn"], c["radii"]
    lmap = {int(e): int(k) for e, k in enumerate(c["layer"])}
    sec = C.circle_section(1.0, 200e3, 0.3)
    Md = E.ULCRBeamAssembler(X, conn, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y,
                             sec.kappa_z).lumped_mass_diag(7.85e-9)
    ctx = C.DeviceContext(X, conn, sec, radii=radii, precision="fp64")
    kcfg = Kc.ContactConfig(mu=c["mu"], k_t_ratio=c["ktr"], exclude_same_layer=True, elem_layer_map=lmap, **kw)
    b = S.DeviceBackend(ctx, M_diag=Md, rho_inf=c["rho_inf"], k_pen=c["k_pen"], smoothing_delta=c["sd"],
                        mu=c["mu"], k_t_ratio=c["ktr"], fixed_dofs=c["fixed"], config=kcfg,
                        device_solve=mode != "host_superlu", core_radii=core,
                        device_control=mode in ("device_ctl", "device_ctl_graph"),
                        use_graph=mode == "device_ctl_graph")
    dev_cc, marks = [], []
    commit, update_geometry = b.commit_coating, b.update_geometry

    def dev_commit():
        dev_cc.append(int((ctx.get_coating()["coating_compression"] > 0).sum()))
        commit()

    def dev_update_geometry(u):   # once per increment, just before its Newton call
        marks.append(b.asm.n_coat_rebuilds)
        update_geometry(u)

    b.commit_coating, b.update_geometry = dev_commit, dev_update_geometry
    res = S.ContactFrictionSolver(b, fixed_dofs=c["fixed"], f_ext_total=c["f"], t_total=c["t_total"],
                                  config=cfg).solve()
    marks.append(b.asm.n_coat_rebuilds)
    assert len(marks) == len(res.calls) + 1
    return res, dev_cc, list(np.diff(marks))


@pytest.mark.parametrize("mode", ["host_superlu", "device_hostctl", "device_ctl", "device_ctl_graph"])
def test_coating_newton_calls_match_oracle(gpu, mode):
    """The solver loop with the coating on (no initial-penetration error, the
    dashpot step per increment, compression committed at every converged
    step) over the device assembly and over the oracle: host SuperLU, device
    attempts with the tests on the host, and device attempts under device
    control (eager and from graphs), where a pair that starts pressing its
    coating within a call stops the attempt at XK_NC_REBUILD and the host
    rebuilds the pattern.  On this squeeze the coating engages where the
    reference's Newton starts to stall (calls that end at 30 attempts, after
    which the two runs' rounding differences grow): the calls are compared up
    to the first converged call that ends with compressed pairs -- same load
    fractions, attempts, convergence, active counts and residual histories --
    and the pattern is rebuilt within those calls as often as by the
    host-tested attempts."""
    ref = coated_squeeze_oracle()
    res, dev_cc, rebuilds = coated_squeeze_device(mode)
    print([(k.load_frac, k.n_attempts, k.n_active, k.converged) for k in res.calls], dev_cc, rebuilds)
    conv = [k for k, call in enumerate(res.calls) if call.converged]
    n_cmp = next(k for j, k in enumerate(conv) if dev_cc[j] > 0) + 1
    assert n_cmp <= len(ref.calls)
    for a, r in zip(res.calls[:n_cmp], ref.calls[:n_cmp]):
        assert (a.load_frac, a.n_attempts, a.n_active, a.converged) == (r.load_frac, r.n_attempts, r.n_active,
                                                                        r.converged)
        np.testing.assert_allclose(a.res_history, r.res_history, rtol=1e-4, atol=0)
    assert sum(rebuilds[:n_cmp]) >= 1, "no pair started pressing its coating within a call"
    if mode in ("device_ctl", "device_ctl_graph"):
        _, _, host_rebuilds = coated_squeeze_device("device_hostctl")
        assert rebuilds[:n_cmp] == host_rebuilds[:n_cmp]
