"""Write the golden fixtures of the assembly path from the reference itself.

Runs ONLY in the build container, where the reference checkout is mounted at
/root/reference (read-only; bytecode writing is disabled).  It imports the
reference's own classes, feeds them inputs made by this repository's strand
generator, and stores inputs and outputs as compressed ``.npz`` data next to
this script.  The fixtures pin what closed forms leave open at the 1e-10 bar:
the exact algebra of the symmetrised CR analytical tangent, the friction
tangents (sign convention of the material block, geometric block), the
friction-loop state quirks (stale p_n, status rewrite, z_t overwrite) and the
composed R / K_T of whole Newton attempts across two increments, and the
trajectory of the reference's own dynamic Newton solver (attempt counts,
cutbacks, active counts, residual histories) on a small strand, which only
running it can pin, and the strand generator's layouts (angle offsets, lay
radii, numbering) over its branches.

Usage:  PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py
        [element|assembly|newton|newton_friction|newton_91|strategy|dynamics|mesh ...]
"""

from __future__ import annotations

import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"

sys.dont_write_bytecode = True
sys.path.insert(0, REPO)
sys.path.insert(0, REF)

mesh_mod = importlib.import_module("xkep-cae_amd.mesh")

from xkep_cae.contact._contact_pair import _ContactConfigInput, _ContactManagerInput  # noqa: E402
from xkep_cae.contact._manager_process import (  # noqa: E402
    DetectCandidatesInput,
    DetectCandidatesProcess,
    UpdateGeometryInput,
    UpdateGeometryProcess,
)
from xkep_cae.contact.contact_force.strategy import HuberContactForceProcess  # noqa: E402
from xkep_cae.contact.friction.strategy import CoulombReturnMappingProcess  # noqa: E402
from xkep_cae.contact.solver._newton_steps import (  # noqa: E402
    ContactForceAssemblyInput,
    ContactForceAssemblyProcess,
    TangentAssemblyInput,
    TangentAssemblyProcess,
)
from xkep_cae.elements._beam_assembler import ULCRBeamAssembler  # noqa: E402
from xkep_cae.elements._beam_cr import (  # noqa: E402
    timo_beam3d_cr_internal_force,
    timo_beam3d_cr_tangent_analytical,
)
from xkep_cae.time_integration.strategy import GeneralizedAlphaProcess  # noqa: E402
from xkep_cae.core.diagnostics import ProcessExecutionLog  # noqa: E402

# The reference writes a usage report under its own docs/ at interpreter exit;
# the checkout must stay untouched, so that report is switched off here.
ProcessExecutionLog.instance()._report_on_exit = False

E, NU, RHO = 200e3, 0.3, 7.85e-9


def section(r):
    d = 2 * r
    A = np.pi * r * r
    Iv = np.pi * d**4 / 64.0
    J = np.pi * d**4 / 32.0
    kap = 6.0 * (1.0 + NU) / (7.0 + 6.0 * NU)
    return dict(E=E, G=E / (2 * (1 + NU)), A=A, Iy=Iv, Iz=Iv, J=J, kappa_y=kap, kappa_z=kap)


def csr_arrays(prefix, K):
    K = K.tocsr()
    K.sort_indices()
    return {f"{prefix}_data": K.data, f"{prefix}_indices": K.indices.astype(np.int32),
            f"{prefix}_indptr": K.indptr.astype(np.int64), f"{prefix}_shape": np.array(K.shape)}


def pair_arrays(prefix, pairs):
    key = np.array([(min(p.elem_a, p.elem_b), max(p.elem_a, p.elem_b)) for p in pairs], np.int64)
    order = np.lexsort((key[:, 1], key[:, 0]))
    P = [pairs[i] for i in order]
    st = [p.state for p in P]
    return {
        f"{prefix}_key": key[order],
        f"{prefix}_s": np.array([s.s for s in st]), f"{prefix}_t": np.array([s.t for s in st]),
        f"{prefix}_gap": np.array([s.gap for s in st]),
        f"{prefix}_normal": np.array([s.normal for s in st]).reshape(-1, 3),
        f"{prefix}_tangent1": np.array([s.tangent1 for s in st]).reshape(-1, 3),
        f"{prefix}_tangent2": np.array([s.tangent2 for s in st]).reshape(-1, 3),
        f"{prefix}_p_n": np.array([s.p_n for s in st]),
        f"{prefix}_z_t": np.array([s.z_t for s in st]).reshape(-1, 2),
        f"{prefix}_k_pen": np.array([s.k_pen for s in st]),
        f"{prefix}_k_t": np.array([s.k_t for s in st]),
        f"{prefix}_q_trial_norm": np.array([s.q_trial_norm for s in st]),
        f"{prefix}_status": np.array([s.status.value for s in st], np.int8),
        f"{prefix}_stick": np.array([bool(s.stick) for s in st]),
        f"{prefix}_dissipation": np.array([s.dissipation for s in st]),
    }


def element_fixture():
    """CR element f_int / K on a small 7-wire strand at three rotation scales, plus
    the reference's scalar element routines on a few elements."""
    m = mesh_mod.make_strand(7, 1.0, 12.0, 12, lay_angle_deg=10.0)
    sec = section(1.0)
    asm = ULCRBeamAssembler(m.node_coords, m.connectivity, sec["E"], sec["G"], sec["A"],
                            sec["Iy"], sec["Iz"], sec["J"], sec["kappa_y"], sec["kappa_z"])
    out = {"X": m.node_coords, "conn": m.connectivity.astype(np.int64), "radius": np.array(1.0)}
    rng = np.random.default_rng(2024)
    ndof = 6 * m.n_nodes
    scales = [(1e-3, 1e-3), (5e-2, 0.3), (0.2, 1.5)]
    for k, (st, sr) in enumerate(scales):
        u = np.zeros(ndof)
        u.reshape(-1, 6)[:, :3] = rng.normal(0, st, (m.n_nodes, 3))
        u.reshape(-1, 6)[:, 3:] = rng.normal(0, sr, (m.n_nodes, 3))
        out[f"u{k}"] = u
        out[f"f{k}"] = asm.assemble_internal_force(u)
        out.update(csr_arrays(f"K{k}", asm.assemble_tangent(u)))
    # scalar routines (Shepperd trace-first branch order) on elements 0..5 at scale 2
    u = out["u2"]
    fs, Ks = [], []
    for e in range(6):
        n1, n2 = m.connectivity[e]
        ed = np.r_[6 * n1 + np.arange(6), 6 * n2 + np.arange(6)]
        args = (m.node_coords[[n1, n2]], u[ed], sec["E"], sec["G"], sec["A"], sec["Iy"], sec["Iz"],
                sec["J"], sec["kappa_y"], sec["kappa_z"])
        fs.append(timo_beam3d_cr_internal_force(*args))
        Ks.append(timo_beam3d_cr_tangent_analytical(*args))
    out["scalar_f"] = np.array(fs)
    out["scalar_K"] = np.array(Ks)
    out["M_lumped"] = asm.assemble_mass(RHO, lumped=True).diagonal()
    np.savez_compressed(os.path.join(HERE, "elem_small.npz"), **out)


def assembly_fixture():
    """Two increments x two Newton attempts of a 7-wire strand with contact and
    friction, through the reference's own step Processes."""
    r, L, ne = 1.0, 20.0, 20
    m = mesh_mod.make_strand(7, r, L, ne, lay_angle_deg=10.0)
    sec = section(r)
    ndof = 6 * m.n_nodes
    k_pen = 1e4 * E * r
    sdelta = 5000.0 / r
    mu = 0.3
    rho_inf = 0.8
    dt = 1e-5
    asm = ULCRBeamAssembler(m.node_coords, m.connectivity, sec["E"], sec["G"], sec["A"],
                            sec["Iy"], sec["Iz"], sec["J"], sec["kappa_y"], sec["kappa_z"])
    M = asm.assemble_mass(RHO, lumped=True)
    ga = GeneralizedAlphaProcess(M, rho_inf=rho_inf)
    huber = HuberContactForceProcess(ndof, smoothing_delta=sdelta)
    fric = CoulombReturnMappingProcess(ndof, k_pen=k_pen, k_t_ratio=0.5)
    manager = _ContactManagerInput(config=_ContactConfigInput(mu=mu))
    radii = m.radii
    out = {"X": m.node_coords, "conn": m.connectivity.astype(np.int64), "radius": np.array(r),
           "params": np.array([k_pen, sdelta, mu, rho_inf, dt, RHO, 0.5])}
    d0 = DetectCandidatesProcess().process(DetectCandidatesInput(
        manager=manager, node_coords=m.node_coords, connectivity=m.connectivity, radii=radii))
    manager = d0.manager
    out["cand0"] = np.array(sorted(d0.candidates), np.int64).reshape(-1, 2)
    rng = np.random.default_rng(77)
    nn = ne + 1
    fixed = ((np.arange(7) * nn)[:, None] * 6 + np.arange(6)[None]).ravel()
    f_ext = np.zeros(ndof)
    f_ext[(np.arange(7) * nn + nn - 1) * 6 + 4] = 5.0
    out["fixed"] = fixed
    out["f_ext"] = f_ext
    u_ref = np.zeros(ndof)
    u = np.zeros(ndof)
    for inc in range(2):
        u_pred = ga.predict(u, dt)
        base = u_pred + rng.normal(0, 2e-3 * r, ndof)
        base[fixed] = 0.0
        coords = m.node_coords + base.reshape(-1, 6)[:, :3]
        dc = DetectCandidatesProcess().process(DetectCandidatesInput(
            manager=manager, node_coords=coords, connectivity=m.connectivity, radii=radii))
        manager = dc.manager
        out[f"inc{inc}_cand"] = np.array(sorted(dc.candidates), np.int64).reshape(-1, 2)
        ug = UpdateGeometryProcess().process(UpdateGeometryInput(manager=manager, node_coords=coords))
        manager = ug.manager
        out[f"inc{inc}_u_pred"] = u_pred.copy()
        out.update(pair_arrays(f"inc{inc}_start", manager.pairs))
        for it in range(2):
            u = base + (rng.normal(0, 5e-4 * r, ndof) if it else 0.0)
            u[fixed] = 0.0
            out[f"inc{inc}_it{it}_u"] = u.copy()
            fo = ContactForceAssemblyProcess().process(ContactForceAssemblyInput(
                u=u, f_ext=f_ext, fixed_dofs=fixed, manager=manager, node_coords_ref=m.node_coords,
                contact_force_strategy=huber, friction_strategy=fric, coating_strategy=None,
                k_pen=k_pen, mu=mu, u_ref=u_ref, load_frac=1.0, load_frac_prev=0.0,
                increment_display=inc + 1, ndof_per_node=6, use_coating=False,
                assemble_internal_force=asm.assemble_internal_force))
            R = fo.R_u
            ga.correct(u, np.zeros_like(u), dt)
            R = ga.effective_residual(R, dt)
            R[fixed] = 0.0
            to = TangentAssemblyProcess().process(TangentAssemblyInput(
                u=u, manager=manager, contact_force_strategy=huber, friction_strategy=fric,
                coating_strategy=None, assemble_tangent=asm.assemble_tangent, k_pen=k_pen, mu=mu,
                ndof=ndof, coords_def=fo.coords_def, load_frac=1.0, load_frac_prev=0.0,
                use_coating=False))
            K = ga.effective_stiffness(to.K_T, dt)
            p = f"inc{inc}_it{it}"
            out[f"{p}_R"] = R
            out[f"{p}_fc"] = fo.f_c
            out[f"{p}_n_active"] = np.array(sum(1 for q in manager.pairs if q.state.p_n > 0.0))
            out.update(csr_arrays(f"{p}_K", K))
            out.update(pair_arrays(f"{p}_pairs", manager.pairs))
        u_ref = u.copy()
    np.savez_compressed(os.path.join(HERE, "assembly_7wire.npz"), **out)


def newton_case(mu=0.0, force=100.0, nw=7, r=1.0, L=20.0, m=20, ang=10.0):
    """Inputs of the Newton-trajectory fixtures: 7-wire strand (r = 1, L = 20, 20
    elements per wire, lay angle 10 deg), core and outer layer in contact only
    (same-layer exclusion by layer index), first node of every wire clamped,
    radial inward force at the mid node of every outer wire, lumped mass,
    Generalized-alpha rho_inf = 0.8, friction coefficient mu (0 = frictionless),
    penalty 0.5 c0 m_ii (the dynamic estimate of
    contact/penalty/strategy.py:216-262 at the nominal step), increments capped
    at a quarter of the total time."""
    mesh = mesh_mod.make_strand(nw, r, L, m, lay_angle_deg=ang)
    X, conn = mesh.node_coords, mesh.connectivity.astype(np.int64)
    s = section(r)
    ndof = 6 * len(X)
    nn = m + 1
    first = np.arange(nw) * nn
    fixed = (first[:, None] * 6 + np.arange(6)[None, :]).ravel()
    mid = first[1:] + nn // 2
    nrm = X[mid, :2] / np.linalg.norm(X[mid, :2], axis=1)[:, None]
    f_ext = np.zeros(ndof)
    f_ext[mid * 6 + 0] = -force * nrm[:, 0]
    f_ext[mid * 6 + 1] = -force * nrm[:, 1]
    # t_total = 4 T1/40 of one wire (first bending mode, numerical_tests/three_point_bend_jig.py:841-858)
    Iv = s["Iy"]
    f1 = (np.pi / (2.0 * L**2)) * np.sqrt(E * Iv / (RHO * s["A"]))
    t_total = 4.0 * (1.0 / f1 / 40.0)
    rho_inf, dt_max_fraction = 0.8, 0.25
    am = (2 * rho_inf - 1) / (rho_inf + 1)
    af = rho_inf / (rho_inf + 1)
    beta = 0.25 * (1 - am + af) ** 2
    dt_nom = t_total * dt_max_fraction
    k_pen = 0.5 * (1.0 / (beta * dt_nom**2)) * RHO * s["A"] * (L / m) / 2.0
    return dict(mesh=mesh, X=X, conn=conn, s=s, fixed=fixed, f_ext=f_ext, t_total=t_total, rho_inf=rho_inf,
                dt_max_fraction=dt_max_fraction, k_pen=k_pen, mu=mu, smoothing_delta=5000.0 / r, tol_force=1e-6,
                tol_disp=1e-8, max_nr_attempts=30, radius=r, layer=mesh.elem_layer(), nw=nw)


def newton_fixture(mu=0.0, force=100.0, name="newton_7wire.npz", **strand):
    """The reference's own dynamic solver (ContactFrictionProcess) on newton_case():
    every Newton call of every increment attempt (cutbacks included) with its
    time window, attempt count, active count, convergence flags and residual
    history, plus the final displacement."""
    from xkep_cae.contact.solver import _newton_dynamic as nd
    from xkep_cae.contact.solver.process import ContactFrictionProcess
    from xkep_cae.core.data import (AssembleCallbacks, BoundaryData, ContactFrictionInputData,
                                    ContactSetupData, MeshData)

    c = newton_case(mu, force, **strand)
    s = c["s"]
    X, conn = c["X"], c["conn"]
    asm = ULCRBeamAssembler(X, conn, s["E"], s["G"], s["A"], s["Iy"], s["Iz"], s["J"], s["kappa_y"], s["kappa_z"])
    M = asm.assemble_mass(RHO, lumped=True)
    lay = c["layer"]
    cfg = _ContactConfigInput(smoothing_delta=c["smoothing_delta"], exclude_same_layer=True,
                              elem_layer_map={int(e): int(k) for e, k in enumerate(lay)},
                              adjust_initial_penetration=False, dt_max_fraction=c["dt_max_fraction"],
                              k_t_ratio=0.5)
    setup = ContactSetupData(manager=_ContactManagerInput(config=cfg), k_pen=c["k_pen"], mu=c["mu"])
    inp = ContactFrictionInputData(
        mesh=MeshData(node_coords=X, connectivity=conn, radii=np.full(len(conn), c["radius"]), n_strands=c["nw"],
                      layer_ids=lay),
        boundary=BoundaryData(fixed_dofs=c["fixed"], f_ext_total=c["f_ext"]),
        contact=setup,
        callbacks=AssembleCallbacks(assemble_tangent=asm.assemble_tangent,
                                    assemble_internal_force=asm.assemble_internal_force, ul_assembler=asm),
        mass_matrix=M, dt_physical=c["t_total"], rho_inf=c["rho_inf"], tol_force=c["tol_force"],
        tol_disp=c["tol_disp"], max_nr_attempts=c["max_nr_attempts"])
    calls = []
    orig = nd.NewtonDynamicProcess.process

    def logged(self, step_input):
        out = orig(self, step_input)
        calls.append((step_input.load_frac, step_input.load_frac_prev, step_input.dt_sub, out.converged,
                      out.n_attempts, out.n_active, out.diverged, list(out.diagnostics.res_history),
                      list(out.diagnostics.n_active_history)))
        return out

    nd.NewtonDynamicProcess.process = logged
    try:
        res = ContactFrictionProcess().process(inp)
    finally:
        nd.NewtonDynamicProcess.process = orig
    n = len(calls)
    hist_len = np.array([len(cl[7]) for cl in calls])
    out = dict(X=X, conn=conn, radii=np.full(len(conn), c["radius"]), layer=lay, fixed=c["fixed"],
               f_ext=c["f_ext"], M_diag=M.diagonal(),
               params=np.array([c["k_pen"], c["mu"], c["t_total"], c["rho_inf"], c["dt_max_fraction"],
                                c["smoothing_delta"], c["tol_force"], c["tol_disp"], c["max_nr_attempts"], 0.5]),
               call_load_frac=np.array([cl[0] for cl in calls]), call_load_frac_prev=np.array([cl[1] for cl in calls]),
               call_dt=np.array([cl[2] for cl in calls]), call_converged=np.array([cl[3] for cl in calls]),
               call_attempts=np.array([cl[4] for cl in calls]), call_active=np.array([cl[5] for cl in calls]),
               call_diverged=np.array([cl[6] for cl in calls]), call_hist_len=hist_len,
               res_history=np.concatenate([np.asarray(cl[7], float) for cl in calls]) if n else np.zeros(0),
               active_history=np.concatenate([np.asarray(cl[8], np.int64) for cl in calls]) if n else np.zeros(0),
               converged=np.array(res.converged), n_increments=np.array(res.n_increments),
               n_cutbacks=np.array(res.n_cutbacks), u_final=res.u)
    np.savez_compressed(os.path.join(HERE, name), **out)


def strategy_fixture():
    """The reference's ContactGeometryStrategy classes (contact/geometry/strategy.py)
    on a perturbed 7-wire strand: ``detect`` (shared-node filter on and off),
    ``update_geometry`` at a second configuration with a config (g_on 0,
    g_off 1e-6) and without one, a coating config, and
    ``build_constraint_jacobian`` of each result; the three classes of the
    factory give the same pairs.  Pairs are stored sorted by (min, max) element
    key; each Jacobian row is stored with the key of its pair."""
    from xkep_cae.contact.geometry.strategy import _create_contact_geometry_strategy

    m = mesh_mod.make_strand(7, 1.0, 20.0, 20, lay_angle_deg=10.0)
    X, conn = m.node_coords, m.connectivity.astype(np.int64)
    rng = np.random.default_rng(31)
    x0 = X + rng.normal(0.0, 0.02, X.shape)
    x1 = x0 + rng.normal(0.0, 0.01, X.shape)
    ndof = 6 * len(X)
    out = {"X": X, "conn": conn, "radii": m.radii, "x0": x0, "x1": x1}

    def jac(prefix, strat, pairs):
        G, act = strat.build_constraint_jacobian(pairs, ndof, 6)
        G = G.tocsr()
        keys = np.array([(min(pairs[i].elem_a, pairs[i].elem_b), max(pairs[i].elem_a, pairs[i].elem_b))
                         for i in act], np.int64).reshape(-1, 2)
        out[f"{prefix}_G_keys"] = keys
        out.update(csr_arrays(f"{prefix}_G", G))

    for mode in ("point_to_point", "line_to_line", "mortar"):
        strat = _create_contact_geometry_strategy(mode=mode)
        pairs = strat.detect(x0, conn, m.radii)
        if mode != "point_to_point":
            ref = out["det_key"]
            key = pair_arrays("tmp", pairs)["tmp_key"]
            assert np.array_equal(key, ref), mode
            continue
        out.update(pair_arrays("det", pairs))
        jac("det", strat, pairs)
        out["det_gap_api"] = np.array([strat.compute_gap(p, x0) for p in pairs])
        out["det_gap_api_key"] = np.array([(min(p.elem_a, p.elem_b), max(p.elem_a, p.elem_b)) for p in pairs])
        cfg = _ContactConfigInput(g_on=0.0, g_off=1e-6)
        strat.update_geometry(pairs, x1, config=cfg)
        out.update(pair_arrays("upd", pairs))
        jac("upd", strat, pairs)
        strat.update_geometry(pairs, x0)
        out.update(pair_arrays("upd2", pairs))
        jac("upd2", strat, pairs)
        ccfg = _ContactConfigInput(g_on=0.0, g_off=1e-6, coating_stiffness=1e3)
        strat.update_geometry(pairs, x1, config=ccfg)
        out.update(pair_arrays("coat", pairs))
        out["coat_compression"] = np.array([pairs[i].state.coating_compression for i in
                                            np.lexsort(([max(p.elem_a, p.elem_b) for p in pairs],
                                                        [min(p.elem_a, p.elem_b) for p in pairs]))])
    strat = _create_contact_geometry_strategy(mode="point_to_point", exclude_same_layer=False)
    pairs = strat.detect(x0, conn, m.radii)
    out.update(pair_arrays("all", pairs))
    jac("all", strat, pairs)
    np.savez_compressed(os.path.join(HERE, "strategy_7wire.npz"), **out)


def dynamics_fixture():
    """The reference's consistent mass (``assemble_mass(lumped=False)``, with the
    automatic reference axis and with v_ref = (0, 1, 0)) on a small 7-wire strand,
    and two Generalized-alpha steps with that mass and a Rayleigh damping matrix
    C = a M + b K(0): predict, correct, effective residual and stiffness."""
    m = mesh_mod.make_strand(7, 1.0, 12.0, 8, lay_angle_deg=10.0)
    sec = section(1.0)
    args = (m.node_coords, m.connectivity, sec["E"], sec["G"], sec["A"], sec["Iy"], sec["Iz"], sec["J"],
            sec["kappa_y"], sec["kappa_z"])
    asm = ULCRBeamAssembler(*args)
    out = {"X": m.node_coords, "conn": m.connectivity.astype(np.int64)}
    M = asm.assemble_mass(RHO, lumped=False)
    out.update(csr_arrays("M", M))
    out.update(csr_arrays("Mv", ULCRBeamAssembler(*args, v_ref=np.array([0.0, 1.0, 0.0])).assemble_mass(
        RHO, lumped=False)))
    ndof = 6 * m.n_nodes
    K0 = asm.assemble_tangent(np.zeros(ndof))
    a, b = 30.0, 2e-5
    C = (a * M + b * K0).tocsr()
    out.update(csr_arrays("C", C))
    out.update(csr_arrays("K0", K0))
    ga = GeneralizedAlphaProcess(M, damping_matrix=C, rho_inf=0.8)
    rng = np.random.default_rng(41)
    dt = 2e-5
    u = np.zeros(ndof)
    for k in range(2):
        up = ga.predict(u, dt)
        u = up + rng.normal(0.0, 1e-4, ndof)
        R = rng.normal(0.0, 1.0, ndof)
        ga.correct(u, np.zeros(ndof), dt)
        out[f"s{k}_u_pred"] = up
        out[f"s{k}_u"] = u.copy()
        out[f"s{k}_R"] = R
        out[f"s{k}_R_eff"] = ga.effective_residual(R, dt)
        out.update(csr_arrays(f"s{k}_K_eff", ga.effective_stiffness(K0, dt)))
        out[f"s{k}_vel"] = ga.vel.copy()
        out[f"s{k}_acc"] = ga.acc.copy()
    out["params"] = np.array([RHO, 0.8, dt, a, b])
    np.savez_compressed(os.path.join(HERE, "dynamics_7wire.npz"), **out)


# strand generator configurations: (n_strands, wire_diameter, pitch, length, n_elems, keyword arguments)
MESH_CASES = [
    (7, 2.0, 40.0, 40.0, 40, {}),
    (3, 1.0, 30.0, 30.0, 30, {"lay_direction": -1}),
    (19, 1.0, 25.0, 0.0, 40, {"n_pitches": 1.5, "gap": 0.05}),
    (37, 1.0, 60.0, 60.0, 64, {"strand_diameter": 8.2}),
    (10, 1.0, 30.0, 30.0, 32, {"strand_diameter": 5.6, "coating_thickness": 0.05}),
    (13, 0.8, 30.0, 30.0, 32, {"coating_thickness": 0.1, "gap": 0.02}),
    (3, 1.0, 30.0, 30.0, 30, {"strand_diameter": 2.5}),
]


def mesh_fixture():
    """The reference's strand generator on configurations that cover its
    branches (three wires, partial layers, explicit outer diameter with the
    surplus spread over the layers, coating thickness, pitches, lay direction)."""
    import warnings

    from xkep_cae.mesh._twisted_wire import _make_twisted_wire_mesh

    out = {}
    for k, (n, d, pitch, length, ne, kw) in enumerate(MESH_CASES):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            m = _make_twisted_wire_mesh(n, d, pitch, length, ne, **kw)
        out[f"c{k}_X"] = m.node_coords
        out[f"c{k}_conn"] = m.connectivity.astype(np.int64)
        out[f"c{k}_info"] = np.array([[i.strand_id, i.layer, i.angle_offset, i.lay_radius, i.wire_radius,
                                       i.lay_direction] for i in m.strand_infos])
        out[f"c{k}_length"] = np.array(m.length)
    np.savez_compressed(os.path.join(HERE, "mesh_layouts.npz"), **out)


def newton_friction_fixtures():
    """The same solver and strand with Coulomb friction at mu = 0.3 (the
    friction coefficient of the strand workloads), under two squeeze loads.
    Once the friction loop engages, the reference's Newton does not converge
    on this strand (its friction material block enters K_T with the opposite
    sign of the consistent derivative), so these trajectories pin the
    frictional attempts, the divergence detection, the cutbacks and the final
    failure: load 60 gives 10 calls (4 cutbacks, friction engaged in 5 of
    them), load 100 gives 5 calls (one divergence abort)."""
    newton_fixture(0.3, 60.0, "newton_7wire_mu03_f60.npz")
    newton_fixture(0.3, 100.0, "newton_7wire_mu03_f100.npz")


def newton_91wire_fixture():
    """The reference's dynamic solver on the 91-wire strand geometry of the
    benchmark (r = 0.4, L = 80, 150 elements per wire, lay angle 12 deg),
    squeezed by 0.1 per outer wire, Coulomb mu = 0.3: 4 Newton calls, 112
    attempts, friction engaged on 432 pairs in the last two.  Too slow to run
    in this container: the first Newton attempt had not finished after ten
    minutes (the reference's per-pair Python loops over about 6e4
    candidates); make_oracle_trajectory.py writes the same case's trajectory
    from the product's driver on the oracle assembly instead."""
    newton_fixture(0.3, 0.1, "newton_91wire_mu03.npz", nw=91, r=0.4, L=80.0, m=150, ang=12.0)


FIXTURES = {"element": element_fixture, "assembly": assembly_fixture, "newton": newton_fixture,
            "newton_friction": newton_friction_fixtures, "newton_91": newton_91wire_fixture,
            "strategy": strategy_fixture, "dynamics": dynamics_fixture, "mesh": mesh_fixture}

if __name__ == "__main__":
    for name in sys.argv[1:] or list(FIXTURES):
        FIXTURES[name]()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
