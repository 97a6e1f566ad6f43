"""Autotuner of launch shapes and grid cell: every shape gives the same
residual and tangent (bit for bit) and the same active count; the tuned
settings are cached per workload shape."""

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def setup(precision):
    wl, C, A, Kc = pkg("workloads"), pkg("context"), pkg("assembly"), pkg("contact")
    case = wl.C2
    st = wl.make_state(case)
    m = st.mesh
    ctx = C.DeviceContext(m.node_coords, m.connectivity, C.circle_section(case.wire_radius, case.E, case.nu),
                          radii=m.radii, precision=precision)
    cfg = Kc.ContactConfig(mu=case.mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(st.elem_layer))
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu, f_ext=st.f_ext,
                           fixed_dofs=st.fixed_dofs, config=cfg)
    asm.begin_increment(np.zeros(ctx.ndof))
    asm.begin_increment(st.u)
    return ctx, asm


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_launch_shape_gives_the_same_assembly(gpu, precision):
    ctx, asm = setup(precision)
    # every assembly advances the friction history (z_t <- q), so each one
    # starts from the same pair state
    ctx.pairs_snapshot()
    R0, K0, r0, n0 = (t.clone() for t in asm.assemble())
    for pb in (64, 128, 256, 512):
        for gb in (128, 256, 512):
            ctx.set_tuning(pair_block=pb, gather_block=gb)
            ctx.pairs_restore()
            R, K, r2, n = asm.assemble()
            assert gpu.equal(R, R0) and gpu.equal(K, K0) and int(n.item()) == int(n0.item()), (pb, gb)
            assert float(r2.item()) == pytest.approx(float(r0.item()), rel=1e-12)
    with pytest.raises(RuntimeError):
        ctx.set_tuning(pair_block=96)


def test_autotune_picks_valid_settings_and_caches_them(gpu, tmp_path, monkeypatch):
    T = pkg("autotune")
    monkeypatch.setenv("XKAMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    ctx, asm = setup("mixed")
    cand0 = ctx.last_candidates().copy()
    out = T.autotune(asm, calls=3)
    assert not out["cached"]
    assert out["pair_block"] in T.PAIR_BLOCKS and out["gather_block"] in T.GATHER_BLOCKS
    assert out["cell_factor"] in T.CELL_FACTORS
    assert np.array_equal(ctx.last_candidates(), cand0)   # the grid cell never changes the candidate set
    again = T.autotune(asm)
    assert again["cached"] and again["pair_block"] == out["pair_block"]
    assert ctx.get_tuning()["gather_block"] == out["gather_block"]
