"""The uneven batched workload: 64 strand instances (7/19/37/61 wires, 50-150
elements per wire, own lay angle, friction coefficient and end moment),
216,830 elements, assembled by one launch sequence.  Every instance's slice of
the batched residual and tangent equals that instance assembled alone, bit for
bit, with no coupling between instances; small instances also against the
oracle."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import pkg
from oracle import assembly as oasm
from oracle import beam as obeam
from oracle import contact as ocon

pytestmark = pytest.mark.gpu


def assemble(X, conn, radii, layer, u, f_ext, fixed, precision, *, mu=None, mu_elem=None):
    wl, C, A, Kc = pkg("workloads"), pkg("context"), pkg("assembly"), pkg("contact")
    r = float(radii[0])
    E, nu = 200e3, 0.3
    ctx = C.DeviceContext(X, conn, C.circle_section(r, E, nu), radii=radii, precision=precision)
    if mu_elem is not None:
        ctx.set_friction(mu_elem)
        mu = 1.0
    cfg = Kc.ContactConfig(mu=mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(layer))
    asm = A.StrandAssembly(ctx, k_pen=1e4 * E * r, smoothing_delta=5000.0 / r, mu=mu, f_ext=f_ext, fixed_dofs=fixed,
                           config=cfg)
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.begin_increment(u)
    R, K, _, n = asm.assemble()
    return R.cpu().numpy(), asm.K_csr(), int(n.item()), ctx


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_slices_equal_single_instances(gpu, precision):
    wl = pkg("workloads")
    cases, states, b, mu = wl.uneven_batch()
    assert b.connectivity.shape[0] == 216830
    Rb, Kb, nb, _ = assemble(b.node_coords, b.connectivity, b.radii, b.elem_layer, b.u, b.f_ext, b.fixed_dofs,
                             precision, mu_elem=mu)
    Kb = Kb.tocsr()
    n_sum = 0
    for k, (c, st) in enumerate(zip(cases, states)):
        m = st.mesh
        n0, n1 = b.node_ranges[k]
        d0, d1 = 6 * n0, 6 * n1
        # the instance alone, at its place in the batch (same coordinates bit for bit)
        Rs, Ks, ns, _ = assemble(b.node_coords[n0:n1], m.connectivity, m.radii, st.elem_layer, st.u, st.f_ext,
                                 st.fixed_dofs, precision, mu=c.mu)
        assert np.array_equal(Rb[d0:d1], Rs), k
        rows = Kb[d0:d1]
        blk = rows[:, d0:d1]
        assert rows.nnz == blk.nnz, k                 # no entry couples two instances
        assert (blk != Ks.tocsr()).nnz == 0, k
        n_sum += ns
    assert nb == n_sum


def test_small_batched_instances_match_oracle(gpu):
    wl = pkg("workloads")
    cases, states, b, mu = wl.uneven_batch()
    Rb, Kb, _, _ = assemble(b.node_coords, b.connectivity, b.radii, b.elem_layer, b.u, b.f_ext, b.fixed_dofs,
                            "fp64", mu_elem=mu)
    Kb = Kb.tocsr()
    small = [k for k, c in enumerate(cases) if c.n_wires == 7][:3]
    assert small
    for k in small:
        c, st = cases[k], states[k]
        m = st.mesh
        r = c.wire_radius
        sec = obeam.circle_section(r, 200e3, 0.3)
        cfg = ocon.ContactConfig(mu=c.mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(st.elem_layer))
        prob = oasm.Problem(m.node_coords, m.connectivity, sec, c.k_pen, c.smoothing_delta, 0.5, c.mu, cfg)
        p, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, cfg)
        x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
        p, _ = ocon.detect_candidates(p, x, m.connectivity, m.radii, cfg)
        p = ocon.update_geometry(p, x, cfg)
        ref = oasm.assemble_iteration(prob, p, st.u, st.u_ref, st.f_ext, st.fixed_dofs, None, 0.0)
        d0, d1 = 6 * b.node_ranges[k][0], 6 * b.node_ranges[k][1]
        assert np.linalg.norm(Rb[d0:d1] - ref.R) <= 1e-10 * np.linalg.norm(ref.R)
        assert sp.linalg.norm(Kb[d0:d1, d0:d1] - ref.K) <= 1e-10 * sp.linalg.norm(ref.K)
