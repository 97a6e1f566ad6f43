"""The 1000-wire strand (layers of 6k wires, the last one partial; r = 0.3,
100 elements per wire: 100,000 elements, about 0.37 M candidate pairs):
candidate set identical to the oracle's and independent of the grid cell,
FP64 assembly against the oracle, FP32-mixed against FP64."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import pkg
from oracle import assembly as oasm
from oracle import beam as obeam
from oracle import contact as ocon

pytestmark = pytest.mark.gpu


def test_1000_wire_candidates_and_assembly(gpu):
    wl, C, A, T, E, Kc = (pkg(n) for n in ("workloads", "context", "assembly", "time_integration", "elements",
                                           "contact"))
    case = wl.C5
    st = wl.make_state(case)
    m = st.mesh
    assert (m.n_wires, m.n_elems) == (1000, 100000)
    lmap = wl.layer_map(st.elem_layer)
    s = C.circle_section(case.wire_radius, case.E, case.nu)
    Md = E.ULCRBeamAssembler(m.node_coords, m.connectivity, s.E, s.G, s.A, s.Iy, s.Iz, s.J, s.kappa_y,
                             s.kappa_z).lumped_mass_diag(case.rho)
    dt = case.dt()
    out = {}
    x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
    for prec in ("fp64", "mixed", "fp32"):
        ctx = C.DeviceContext(m.node_coords, m.connectivity, s, radii=m.radii, precision=prec)
        ga = T.GeneralizedAlphaProcess(Md, rho_inf=case.rho_inf)
        ga.predict_device(ctx.torch.zeros(ctx.ndof, dtype=ctx.torch.float64, device=ctx.dev), dt)
        asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                               f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, time=ga, dt=dt,
                               config=Kc.ContactConfig(mu=case.mu, exclude_same_layer=True, elem_layer_map=lmap))
        asm.begin_increment(np.zeros(ctx.ndof))
        asm.begin_increment(st.u)
        cand = ctx.last_candidates()
        R, Kv, _, nact = asm.assemble(st.u)
        out[prec] = (R.cpu().numpy(), asm.K_csr(), int(nact.item()), ctx.get_pairs()["status"].copy(), cand)
        if prec == "fp64":
            xd = ctx.to_dev(x)
            raw = [ctx.broadphase(xd, cell_size=h) for h in (0.0, 0.45, 1.3)]
            assert all(np.array_equal(raw[0], r) for r in raw[1:])
    cfg = ocon.ContactConfig(mu=case.mu, exclude_same_layer=True, elem_layer_map=lmap)
    osec = obeam.circle_section(case.wire_radius, case.E, case.nu)
    prob = oasm.Problem(m.node_coords, m.connectivity, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, cfg)
    op, ocand = ocon.detect_candidates(op, x, m.connectivity, m.radii, cfg)
    op = ocon.update_geometry(op, x, cfg)
    ga = oasm.GeneralizedAlpha(obeam.lumped_mass_diag(m.node_coords, m.connectivity, case.rho, osec), case.rho_inf)
    ga.predict(np.zeros_like(st.u), dt)
    ref = oasm.assemble_iteration(prob, op, st.u, st.u_ref, st.f_ext, st.fixed_dofs, ga, dt)
    R64, K64, n64, s64, c64 = out["fp64"]
    assert np.array_equal(c64, ocand) and len(ocand) > 300000
    assert np.linalg.norm(R64 - ref.R) <= 1e-10 * np.linalg.norm(ref.R)
    assert sp.linalg.norm(K64 - ref.K) <= 1e-10 * sp.linalg.norm(ref.K)
    assert n64 == ref.n_active and np.array_equal(s64, ref.pairs.status)
    for prec in ("mixed", "fp32"):
        Rm, Km, nm, sm, cm = out[prec]
        assert np.array_equal(cm, c64), prec
        assert np.linalg.norm(Rm - R64) <= 1e-6 * np.linalg.norm(R64), prec
        assert nm == n64 and np.array_equal(sm, s64), prec
