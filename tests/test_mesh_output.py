"""The inputs and outputs either side of the path: the strand generator
(xkep-cae_amd/mesh.py) and the result formats (xkep-cae_amd/output.py).

Generator: the reference's own layouts (tests/golden/mesh_layouts.npz, written
by tests/golden/make_golden.py from ``_make_twisted_wire_mesh``) over its
branches -- three wires, partial layers, explicit outer diameter, coating
thickness, pitches, lay direction -- must be reproduced exactly; beside them
geometric invariants: every node of a wire sits at its lay radius, the
minimum outer diameter makes neighbouring wires touch, the checks refuse
what the reference refuses.  Output: the reference's ExportProcess tests
(output/tests/test_export.py) -- files written per format, the summary's
values -- and a round trip of the written data."""

import json
import math
import warnings

import numpy as np
import pytest
import yaml

from conftest import golden, pkg


def cases():
    g = golden("mesh_layouts.npz")
    # same list as tests/golden/make_golden.py
    return g, [
        (7, 2.0, 40.0, 40.0, 40, {}),
        (3, 1.0, 30.0, 30.0, 30, {"lay_direction": -1}),
        (19, 1.0, 25.0, 0.0, 40, {"n_pitches": 1.5, "gap": 0.05}),
        (37, 1.0, 60.0, 60.0, 64, {"strand_diameter": 8.2}),
        (10, 1.0, 30.0, 30.0, 32, {"strand_diameter": 5.6, "coating_thickness": 0.05}),
        (13, 0.8, 30.0, 30.0, 32, {"coating_thickness": 0.1, "gap": 0.02}),
        (3, 1.0, 30.0, 30.0, 30, {"strand_diameter": 2.5}),
    ]


def test_generator_reproduces_the_reference_layouts():
    M = pkg("mesh")
    g, cs = cases()
    for k, (n, d, pitch, length, ne, kw) in enumerate(cs):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            m = M.make_twisted_wire_mesh(n, d, pitch, length, ne, **kw)
        info = np.array([[i.strand_id, i.layer, i.angle_offset, i.lay_radius, i.wire_radius, i.lay_direction]
                         for i in m.strand_infos])
        assert np.array_equal(info, g[f"c{k}_info"]), k
        assert np.array_equal(m.node_coords, g[f"c{k}_X"]), k
        assert np.array_equal(m.connectivity, g[f"c{k}_conn"]), k
        assert m.length == float(g[f"c{k}_length"])


def test_nodes_sit_on_their_lay_radius_and_wires_chain():
    M = pkg("mesh")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = M.make_twisted_wire_mesh(19, 1.0, 25.0, 25.0, 40, gap=0.05)
    for info, (n0, n1), (e0, e1) in zip(m.strand_infos, m.strand_node_ranges, m.strand_elem_ranges):
        xy = np.linalg.norm(m.node_coords[n0:n1, :2], axis=1)
        np.testing.assert_allclose(xy, info.lay_radius, rtol=0, atol=1e-12)
        np.testing.assert_allclose(m.node_coords[n0:n1, 2], np.linspace(0, 25.0, n1 - n0), rtol=0, atol=1e-12)
        assert np.array_equal(m.connectivity[e0:e1], np.stack([np.arange(n0, n1 - 1), np.arange(n0 + 1, n1)], 1))
    # layer k at k (d + gap), gap raised to the chord-sagitta safe gap if smaller
    g_min = M.min_safe_gap(19, 0.5, 25.0, 40, 1.0)
    gap = max(0.05, g_min)
    for info in m.strand_infos:
        assert info.lay_radius == pytest.approx(info.layer * (1.0 + gap), rel=1e-15)


@pytest.mark.parametrize("n", [3, 7, 19, 37])
def test_minimum_outer_diameter_makes_wires_touch(n):
    """At the minimum outer diameter, the binding constraint of every layer
    holds with equality: neighbours in a layer (circumferential) or the layer
    below (radial) at one wire diameter; the outermost wires touch the outer
    circle."""
    M = pkg("mesh")
    d = 1.0
    D = M.minimum_strand_diameter(n, d)
    lay = M.strand_layout(n, d / 2, strand_diameter=D)
    radii = sorted({round(i.lay_radius, 12) for i in lay})
    assert radii[-1] + d / 2 == pytest.approx(D / 2, rel=1e-14)
    counts = {}
    for i in lay:
        counts.setdefault(i.layer, []).append(i)
    prev = 0.0 if 0 in counts else None     # three wires: no core below the layer
    for k in sorted(counts):
        if k == 0:
            continue
        r = counts[k][0].lay_radius
        chord = 2 * r * math.sin(math.pi / len(counts[k]))
        radial = math.inf if prev is None else r - prev
        assert min(chord, radial) >= d - 1e-12                    # no overlap
        assert min(abs(chord - d), abs(radial - d)) < 1e-12        # and touching
        prev = r
    with pytest.raises(ValueError):
        M.strand_layout(n, d / 2, strand_diameter=D * (1 - 1e-6))


def test_generator_refusals():
    M = pkg("mesh")
    with pytest.raises(ValueError):
        M.make_twisted_wire_mesh(7, 1.0, 10.0, 10.0, 15)              # 15 elements per pitch
    with pytest.raises(ValueError):
        M.make_twisted_wire_mesh(7, 1.0, 10.0, 10.0, 40, min_elems_per_pitch=8)
    with pytest.warns(UserWarning):
        M.make_twisted_wire_mesh(7, 1.0, 10.0, 10.0, 20, gap=0.0)     # raised to the safe gap


def test_strand_mesh_process():
    M = pkg("mesh")
    r = M.StrandMeshProcess().process(M.StrandMeshConfig())
    assert (M.StrandMeshConfig().n_strands, M.StrandMeshConfig().wire_radius,
            M.StrandMeshConfig().n_elements_per_pitch) == (7, 0.5, 16)
    assert r.mesh.node_coords.shape == (7 * 17, 3) and r.mesh.connectivity.shape == (7 * 16, 2)
    assert r.mesh.n_strands == 7 and np.array_equal(np.unique(r.mesh.layer_ids), [0, 1])
    with pytest.raises(Exception):
        M.StrandMeshConfig().n_strands = 3                              # frozen


def test_export_process(tmp_path):
    O = pkg("output")
    res = O.SolverResultData(u=np.array([0.1, 0, 0, 0, 0, 0, 0.2, 0, 0, 0, 0, 0]), converged=True,
                             n_increments=5, total_attempts=20, contact_force_history=[1.0, 2.0, 3.0])
    r = O.ExportProcess().process(O.ExportConfig(solver_result=res, mesh=None, output_dir=str(tmp_path / "a")))
    assert len(r.exported_files) == 2
    np.testing.assert_array_equal(np.loadtxt(r.exported_files[0], delimiter=","), res.u)
    np.testing.assert_array_equal(np.loadtxt(r.exported_files[1], delimiter=","), [1.0, 2.0, 3.0])
    assert open(r.exported_files[1]).readline().strip() == "# contact_force_norm"
    r = O.ExportProcess().process(O.ExportConfig(solver_result=res, mesh=None, output_dir=str(tmp_path / "b"),
                                                 formats=("json",)))
    assert len(r.exported_files) == 1
    data = json.load(open(r.exported_files[0]))
    assert data == {"converged": True, "n_increments": 5, "total_attempts": 20, "max_displacement": 0.2,
                    "elapsed_seconds": 0.0}
    r = O.ExportProcess().process(O.ExportConfig(solver_result=res, mesh=None, output_dir=str(tmp_path / "c"),
                                                 formats=("csv", "json")))
    assert len(r.exported_files) == 3


def test_solver_result_and_run_summary(tmp_path):
    """A solve's result through SolverResultData into the YAML run record."""
    from _oracle_backend import OracleBackend
    from oracle import beam as obeam
    from oracle import contact as ocon

    S, O, wl = pkg("solver"), pkg("output"), pkg("workloads")
    c = wl.squeeze_case(20.0)
    osec = obeam.circle_section(1.0, 200e3, 0.3)
    lmap = {int(e): int(k) for e, k in enumerate(c["layer"])}
    ocfg = ocon.ContactConfig(mu=c["mu"], exclude_same_layer=True, elem_layer_map=lmap)
    b = OracleBackend(c["X"], c["conn"], c["radii"], osec,
                      M_diag=obeam.lumped_mass_diag(c["X"], c["conn"], 7.85e-9, osec), rho_inf=c["rho_inf"],
                      k_pen=c["k_pen"], smoothing_delta=c["sd"], mu=c["mu"], k_t_ratio=c["ktr"], cfg=ocfg)
    b.fixed = c["fixed"]
    res = S.ContactFrictionSolver(b, fixed_dofs=c["fixed"], f_ext_total=c["f"], t_total=c["t_total"],
                                  config=S.SolverConfig(**c["cfg"])).solve()
    d = O.solver_result_data(res, elapsed_seconds=1.234)
    assert d.converged and d.load_history[-1] == 1.0 and d.n_increments == res.n_increments
    rec = O.write_run_summary(tmp_path / "results" / "result_x.yaml", "x", d, deflection_mm=0.1234)
    back = yaml.safe_load(open(tmp_path / "results" / "result_x.yaml"))
    assert back == rec
    assert back["elapsed_seconds"] == 1.2 and back["final_frac"] == 1.0 and back["deflection_mm"] == 0.1234
