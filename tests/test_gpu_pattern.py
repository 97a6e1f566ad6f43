"""Tangent pattern of the direct element path built on the device, against
the host build of the same library (``host_pattern=True``) and against the
block pattern stated from its definition in numpy: the union of every
element's 4x4 group blocks and, for every pair that contributes within the
increment (status != INACTIVE or p_n > 0), the 4x4 translational blocks
between its four nodes.  The pattern is this library's own: it holds every
entry the reference's sparse sum of element and pair tangents can produce
within the increment.

Identical patterns and identical contribution lists give bitwise identical
residuals and tangents, so the assembly outputs are compared with ==.
The hand-made pair lists hit the build's special cases: no pairs, no
contributing pair, pairs whose blocks an element already holds (two elements
of one wire), consecutive node numbers that no element joins (the last node
of one wire and the first of the next), and repeated pairs."""

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def expected_pattern(n_nodes, conn, nodes4, contributing):
    """Scalar CSR (indptr, indices) of the block pattern, from its definition."""
    rows = [set() for _ in range(2 * n_nodes)]
    for n1, n2 in conn:
        g = (2 * n1, 2 * n1 + 1, 2 * n2, 2 * n2 + 1)
        for a in g:
            rows[a].update(g)
    for q in np.nonzero(contributing)[0]:
        g = [2 * int(n) for n in nodes4[q]]
        for a in g:
            rows[a].update(g)
    indptr, indices = [0], []
    for gr in range(2 * n_nodes):
        cols = sorted(rows[gr])
        for _ in range(3):
            indices.extend(3 * c + k for c in cols for k in range(3))
            indptr.append(len(indices))
    return np.array(indptr, np.int64), np.array(indices, np.int32)


def pair_nodes(conn, ea, eb):
    return np.stack([conn[ea, 0], conn[ea, 1], conn[eb, 0], conn[eb, 1]], axis=1)


def assert_pattern(ctx, want):
    indptr, indices, nnz = ctx.pattern()
    assert nnz == len(want[1])
    assert np.array_equal(indptr, want[0])
    assert np.array_equal(indices, want[1])


def twins(mesh_x, conn, radii, r, precision):
    C = pkg("context")
    sec = C.circle_section(r, 200e3, 0.3)
    dev = C.DeviceContext(mesh_x, conn, sec, radii=radii, precision=precision)
    host = C.DeviceContext(mesh_x, conn, sec, radii=radii, precision=precision, host_pattern=True)
    assert dev.direct and host.direct
    return dev, host


def outputs(asm):
    R, K, res2, nact = asm.assemble()
    asm.ctx.torch.cuda.synchronize()
    return R.cpu().numpy(), K.cpu().numpy(), int(nact.item())


def assert_same(o1, o2):
    assert np.array_equal(o1[0], o2[0], equal_nan=True)
    assert np.array_equal(o1[1], o2[1], equal_nan=True)
    assert o1[2] == o2[2]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_pattern_equals_host_build_over_increments_at_91_wires(gpu, precision):
    """91-wire strand, same-layer filter: three increments (reference
    configuration, half the deformed state, the deformed state), pattern against the
    definition and the host build, assembly bitwise against the host build."""
    wl = pkg("workloads")
    A, Kc = pkg("assembly"), pkg("contact")
    case = wl.C3
    st = wl.make_state(case)
    m = st.mesh
    cfg = Kc.ContactConfig(mu=case.mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(st.elem_layer))
    asms = [A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                             f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, config=cfg)
            for ctx in twins(m.node_coords, m.connectivity, m.radii, case.wire_radius, precision)]
    n_contrib = []
    for u in (np.zeros_like(st.u), 0.5 * st.u, st.u):
        outs = []
        for asm in asms:
            asm.begin_increment(u)
            d = asm.ctx.get_pairs()
            contributing = (d["status"] != 0) | (d["p_n"] > 0.0)
            if asm is asms[0]:
                n_contrib.append(int(contributing.sum()))
                want = expected_pattern(asm.ctx.n_nodes, m.connectivity,
                                        pair_nodes(m.connectivity, d["elem_a"], d["elem_b"]), contributing)
            assert_pattern(asm.ctx, want)
            outs.append(outputs(asm))
        assert_same(outs[0], outs[1])
    assert n_contrib[-1] > 0 and len(set(n_contrib)) > 1, n_contrib


def test_device_pattern_equals_host_build_on_batched_instances(gpu):
    """Three 37-wire instances in one context (mixed form)."""
    wl = pkg("workloads")
    A = pkg("assembly")
    case = wl.C2
    b = wl.concat_states([wl.make_state(case, seed=s) for s in (37, 38, 39)])
    asms = [A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                             f_ext=b.f_ext, fixed_dofs=b.fixed_dofs)
            for ctx in twins(b.node_coords, b.connectivity, b.radii, case.wire_radius, "mixed")]
    outs = []
    for asm in asms:
        asm.begin_increment(np.zeros_like(b.u))
        asm.begin_increment(b.u)
        outs.append((outputs(asm), asm.ctx.pattern()))
    assert_same(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1][0], outs[1][1][0]) and np.array_equal(outs[0][1][1], outs[1][1][1])


def hand_made_pairs(conn, n_per_wire, rng):
    """Element pairs of the 7-wire mesh that hit the special cases."""
    ne = len(conn)
    wire = np.arange(ne) // n_per_wire
    ea, eb = [], []
    # two elements of one wire: nodes that elements already couple
    for e in (3, 10, 2 * n_per_wire + 5):
        ea += [e, e]
        eb += [e + 2, e + 3]
    # last element of wire w, first element of wire w + 1: node numbers n, n + 1 that no element joins
    for w in range(3):
        ea.append((w + 1) * n_per_wire - 1)
        eb.append((w + 1) * n_per_wire)
    # random pairs across wires, some repeated, some in both orders
    while len(ea) < 60:
        a, b = rng.integers(0, ne, 2)
        if wire[a] != wire[b]:
            ea.append(int(a))
            eb.append(int(b))
    ea += [ea[-1], eb[-2]]
    eb += [eb[-1], ea[-2]]
    return np.array(ea, np.int32), np.array(eb, np.int32)


def pair_state(n, rng, status, p_n):
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    t1 = np.cross(nrm, [0.0, 0.0, 1.0])
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    return {"s": rng.uniform(0.1, 0.9, n), "t": rng.uniform(0.1, 0.9, n), "gap": rng.uniform(-1e-3, 1e-3, n),
            "normal": nrm, "tangent1": t1, "p_n": p_n, "k_pen": np.full(n, 1e3), "k_t": np.full(n, 5e2),
            "status": status.astype(np.int8)}


@pytest.mark.parametrize("kind", ["no_pairs", "none_contributing", "mixed_contributing", "all_contributing"])
def test_device_pattern_on_hand_made_pair_lists(gpu, kind):
    wl = pkg("workloads")
    A = pkg("assembly")
    case = wl.C1
    st = wl.make_state(case)
    m = st.mesh
    conn = np.asarray(m.connectivity)
    rng = np.random.default_rng(11)
    ea, eb = hand_made_pairs(conn, case.n_elems_per_wire, rng)
    n = len(ea)
    if kind == "no_pairs":
        ea, eb, state = ea[:0], eb[:0], None
        contributing = np.zeros(0, bool)
    else:
        status = {"none_contributing": np.zeros(n), "all_contributing": np.ones(n),
                  "mixed_contributing": rng.integers(0, 3, n)}[kind]
        p_n = np.zeros(n) if kind == "none_contributing" else np.where(rng.random(n) < 0.3, 1.0, 0.0)
        state = pair_state(n, rng, status, p_n)
        contributing = (status != 0) | (p_n > 0.0)
    want = expected_pattern(m.node_coords.shape[0], conn, pair_nodes(conn, ea, eb), contributing)
    outs = []
    for ctx in twins(m.node_coords, m.connectivity, m.radii, case.wire_radius, "fp64"):
        asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                               f_ext=st.f_ext, fixed_dofs=st.fixed_dofs)
        asm.set_u(st.u)
        ctx.set_pairs(ea, eb, state)
        assert_pattern(ctx, want)
        outs.append(outputs(asm))
    assert_same(outs[0], outs[1])
