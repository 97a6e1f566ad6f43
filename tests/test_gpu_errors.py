"""Error paths of the device path, against the reference's conventions
(contact/solver/_newton_steps.py:350-366: a failed linear solve is
success=False and ends the Newton call; elements/_beam_cr.py:54-55: a
zero-length element raises ValueError):

- non-finite coordinates are refused by the broadphase before any cell loop;
- a NaN displacement sets XK_FLAG_NONFINITE and ends the Newton call as failed
  (device attempts and host SuperLU attempts alike);
- a zero-length element sets XK_FLAG_DEGENERATE and the drop-ins raise;
- a tangent without stiffness sets XK_FLAG_SINGULAR in the device solve and
  ends the Newton call as failed;
- argument errors of the device Newton control's coating probe and of the
  optional per-attempt GA state come back as errors before any launch."""

import numpy as np
import pytest

from conftest import pkg


def _strand(precision="fp64", n_elems=12):
    wl, C = pkg("workloads"), pkg("context")
    case = wl.StrandCase("err_7wire", 7, 1.0, 12.0, n_elems, 10.0, seed=5)
    st = wl.make_state(case)
    m = st.mesh
    sec = C.circle_section(case.wire_radius, case.E, case.nu)
    ctx = C.DeviceContext(m.node_coords, m.connectivity, sec, radii=m.radii, precision=precision)
    return case, st, ctx


def _backend(case, st, ctx, *, device_solve):
    S, E, C = pkg("solver"), pkg("elements"), pkg("context")
    m = st.mesh
    sec = C.circle_section(case.wire_radius, case.E, case.nu)
    Md = E.ULCRBeamAssembler(m.node_coords, m.connectivity, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J,
                             sec.kappa_y, sec.kappa_z).lumped_mass_diag(case.rho)
    return S.DeviceBackend(ctx, M_diag=Md, rho_inf=0.8, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta,
                           mu=case.mu, fixed_dofs=st.fixed_dofs, device_solve=device_solve)


@pytest.mark.gpu
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_broadphase_refuses_non_finite_coordinates(gpu, bad):
    torch = gpu
    _, st, ctx = _strand()
    x = st.mesh.node_coords.copy()
    x[5, 1] = bad
    print(f"broadphase with a {bad} coordinate", flush=True)
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.detect_candidates(ctx.to_dev(x))
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.broadphase(ctx.to_dev(x))
    # the context stays usable: the finite coordinates give the set they gave before
    n = ctx.detect_candidates(ctx.to_dev(st.mesh.node_coords))
    torch.cuda.synchronize()
    assert n > 0


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["device_control", "device_hostctl", "host_superlu"])
def test_nan_displacement_flags_and_fails_the_call(gpu, mode):
    S = pkg("solver")
    case, st, ctx = _strand()
    device_solve = mode != "host_superlu"
    b = _backend(case, st, ctx, device_solve=device_solve)
    b.device_control = mode == "device_control"
    u0 = np.zeros(ctx.ndof)
    b.detect(u0)
    b.update_geometry(u0)
    u = u0.copy()
    free_rot = 6 * 7 + 4            # a rotation of a free node: only the element kernel sees it
    assert free_rot not in set(st.fixed_dofs.tolist())
    u[free_rot] = np.nan
    print(f"NaN displacement, {mode}", flush=True)
    cfg = S.SolverConfig(max_nr_attempts=10, tol_force=1e-8)
    out = S.newton_dynamic(b, u, u0.copy(), st.f_ext, st.fixed_dofs, 1e-4, 1.0, False, cfg)
    assert not out.converged and out.n_attempts == 1
    assert b.flags & S.FLAG_NONFINITE
    if mode == "device_control":
        assert b.control_reasons == [1]   # XK_NCR_NONFINITE


@pytest.mark.gpu
def test_zero_length_element_raises(gpu):
    C, E, S = pkg("context"), pkg("elements"), pkg("solver")
    X = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 2.0]])
    conn = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    sec = C.circle_section(0.1, 200e3, 0.3)
    asm = E.ULCRBeamAssembler(X, conn, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y, sec.kappa_z)
    print("zero-length element", flush=True)
    with pytest.raises(ValueError, match="zero-length"):
        asm.assemble_internal_force(np.zeros(24))
    with pytest.raises(ValueError, match="zero-length"):
        asm.assemble_tangent(np.zeros(24))
    # a proper chain passes
    X2 = X.copy()
    X2[2:, 2] += 1.0
    asm2 = E.ULCRBeamAssembler(X2, conn, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y, sec.kappa_z)
    assert np.all(np.isfinite(asm2.assemble_internal_force(np.zeros(24))))
    ctx = C.DeviceContext(X, conn, sec, radii=np.full(3, 0.1))
    b = S.DeviceBackend(ctx, M_diag=np.ones(24), rho_inf=0.8, k_pen=1.0, smoothing_delta=0.0, mu=0.0)
    with pytest.raises(ValueError, match="XK_FLAG_DEGENERATE"):
        b.internal_force(np.zeros(24))


@pytest.mark.gpu
def test_singular_tangent_flags_and_fails_the_call(gpu):
    """E = 0 (torsion only), static, nothing fixed, a straight wire along z:
    every translational and bending row of the tangent is exactly zero, the
    chain factorization meets a zero pivot, the solve fails."""
    torch = gpu
    C, S, A = pkg("context"), pkg("solver"), pkg("assembly")
    X = np.array([[0.0, 0.0, float(k)] for k in range(5)])
    conn = np.array([[k, k + 1] for k in range(4)], np.int32)
    sec = C.Section(0.0, 1.0e4, 0.0314, 7.85e-5, 7.85e-5, 1.57e-4, 0.886, 0.886)
    ctx = C.DeviceContext(X, conn, sec, radii=np.full(4, 0.1))
    asm = A.StrandAssembly(ctx, k_pen=1.0, smoothing_delta=0.0, mu=0.0, f_ext=np.ones(30))
    asm.begin_increment(np.zeros(30))
    R, K, _, _ = asm.assemble(np.zeros(30))
    du = torch.empty(30, dtype=torch.float64, device=ctx.dev)
    print("singular tangent: device solve", flush=True)
    ctx.linear_solve(asm.u, asm.fixed_mask, R, du, None)
    torch.cuda.synchronize()
    assert ctx.status_flags() & S.FLAG_SINGULAR
    b = S.DeviceBackend(ctx, M_diag=np.zeros(30), rho_inf=0.8, k_pen=1.0, smoothing_delta=0.0, mu=0.0,
                        device_solve=True)
    b.detect(np.zeros(30))
    b.update_geometry(np.zeros(30))
    u = np.zeros(30)
    print("singular tangent: Newton call", flush=True)
    out = S.newton_dynamic(b, u, np.zeros(30), np.ones(30), np.zeros(0, np.int64), 0.0, 1.0, False,
                           S.SolverConfig(max_nr_attempts=5))
    assert not out.converged and out.n_attempts == 1
    assert b.flags & S.FLAG_SINGULAR
    assert b.control_reasons == [4]   # XK_NCR_SINGULAR: stopped on the device after the solve
    assert np.all(u == 0.0)    # no update after the failed solve


@pytest.mark.gpu
def test_control_probe_and_partial_ga_state_are_refused(gpu):
    """Argument errors come back as errors, before any launch: the coating
    probe of a device-controlled call without an attached control record, and
    an assembly given only one of the two per-attempt GA state pointers."""
    L, T, A = pkg("_lib"), pkg("time_integration"), pkg("assembly")
    case, st, ctx = _strand()
    ctx.set_coating(0.1 * case.k_pen, 0.0, 0.2, 0.5, None)
    with pytest.raises(RuntimeError, match="no Newton control record attached"):
        ctx.coating_probe_control(ctx.to_dev(np.zeros(ctx.ndof)))
    ctx.set_coating(0.0)
    E, C = pkg("elements"), pkg("context")
    m = st.mesh
    sec = C.circle_section(case.wire_radius, case.E, case.nu)
    Md = E.ULCRBeamAssembler(m.node_coords, m.connectivity, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J,
                             sec.kappa_y, sec.kappa_z).lumped_mass_diag(case.rho)
    ga = T.GeneralizedAlphaProcess(Md, rho_inf=0.8)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, time=ga, dt=1e-5)
    asm.begin_increment(np.zeros(ctx.ndof))
    asm._dyn.vel = None   # acc set, vel not
    with pytest.raises(RuntimeError, match="acc and vel are both set or both null"):
        asm.assemble(st.u)
