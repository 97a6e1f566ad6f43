"""Why the FP32-mixed form keeps the element and contact forces, and the pair
geometry that decides them, in FP64.

The reference solver's frictional Newton trajectory
(tests/golden/newton_7wire_mu03_f60.npz, 10 calls with 4 cutbacks) is run by
the product driver on the oracle assembly three times: in FP64; with every
element's internal force rounded to FP32 before the global sum (the least
error an FP32 element force can carry); and with the segment closest-point
search in FP32 (coordinate differences formed in FP64, then rounded, the rest
in FP32).  Both FP32 variants keep this strand's attempt and active counts,
but both move the per-attempt residual ratio ||R|| / f_ref by far more than
the mixed form's 1e-6 bound on the residual error -- near convergence ||R|| is
a small difference of large element and contact forces, so their FP32
rounding dominates it.  (The 91-wire frictional squeeze is more sensitive
still: there the FP32 rounding of the solved tangent alone changes the
trajectory.)  Test infrastructure: imports the oracle."""

import numpy as np

from conftest import pkg
from oracle import beam as obeam
from oracle import contact as ocon

import test_newton_trajectory as T

FIXTURE = "newton_7wire_mu03_f60.npz"


def _elem_force_fp32(orig):
    def f(*a, **k):
        f_e, K_e = orig(*a, **k)
        if f_e is not None:
            f_e = f_e.astype(np.float32).astype(np.float64)
        return f_e, K_e
    return f


def _closest_points_fp32(xA0, xA1, xB0, xB1, tol_parallel=1e-10):
    """ocon.closest_point_segments with the arithmetic in FP32."""
    f = np.float32
    dA, dB, w0 = (xA1 - xA0).astype(f), (xB1 - xB0).astype(f), (xA0 - xB0).astype(f)
    a, b, c = (np.einsum("ij,ij->i", p, q) for p, q in ((dA, dA), (dA, dB), (dB, dB)))
    d, e = np.einsum("ij,ij->i", dA, w0), np.einsum("ij,ij->i", dB, w0)
    det = a * c - b * b
    par = det < f(tol_parallel) * np.maximum(a * c, f(1e-30))
    sdet = np.where(par, f(1), det)
    s_unc = np.where(par, f(0), (b * e - c * d) / sdet)
    t_unc = np.where(par, f(0), (a * e - b * d) / sdet)
    c_ok, a_ok = c > tol_parallel, a > tol_parallel
    sc, sa = np.where(c_ok, c, f(1)), np.where(a_ok, a, f(1))
    s = np.where(par, f(0), np.clip(s_unc, 0, 1))
    t = np.where(par, np.where(c_ok, np.clip(e / sc, 0, 1), 0).astype(f), np.clip(t_unc, 0, 1))
    s_cl, t_cl = (s_unc < 0) | (s_unc > 1), (t_unc < 0) | (t_unc > 1)
    t = np.where(s_cl & ~par, np.where(c_ok, np.clip((b * s + e) / sc, 0, 1), 0), t)
    need = t_cl & ~par
    s = np.where(need, np.where(a_ok, np.clip((b * t - d) / sa, 0, 1), 0), s)
    t = np.where(need, np.where(c_ok, np.clip((b * s + e) / sc, 0, 1), 0), t)
    diff = w0 + s[:, None] * dA - t[:, None] * dB
    dist = np.sqrt(np.einsum("ij,ij->i", diff, diff))
    nrm = diff / np.maximum(dist, f(1e-30))[:, None]
    nrm[dist < 1e-30] = [0, 0, 1]
    return s.astype(float), t.astype(float), dist.astype(float), nrm.astype(float), par


def _solve():
    from _oracle_backend import OracleBackend

    g, prm, cfg = T.case(FIXTURE)
    sec = obeam.circle_section(float(g["radii"][0]), 200e3, 0.3)
    ocfg = ocon.ContactConfig(mu=prm["mu"], exclude_same_layer=True,
                              elem_layer_map={int(e): int(k) for e, k in enumerate(g["layer"])})
    b = OracleBackend(g["X"], g["conn"], g["radii"], sec, M_diag=g["M_diag"], rho_inf=prm["rho_inf"],
                      k_pen=prm["k_pen"], smoothing_delta=prm["smoothing_delta"], mu=prm["mu"],
                      k_t_ratio=prm["k_t_ratio"], cfg=ocfg)
    b.fixed = g["fixed"]
    res = pkg("solver").ContactFrictionSolver(b, fixed_dofs=g["fixed"], f_ext_total=g["f_ext"],
                                               t_total=prm["t_total"], config=cfg).solve()
    return g, res


def _max_dev(res, ref):
    return max(float(np.max(np.abs(np.array(c.res_history) - np.array(r.res_history)) / np.array(r.res_history)))
               for c, r in zip(res.calls, ref.calls))


def test_fp32_forces_or_pair_geometry_exceed_the_mixed_residual_bound(monkeypatch):
    g, ref = _solve()
    T.check_trajectory(g, ref, "oracle")
    devs = {}
    for name, mod, attr, repl in (("element force", obeam, "cr_beam_element_batch",
                                   _elem_force_fp32(obeam.cr_beam_element_batch)),
                                  ("pair geometry", ocon, "closest_point_segments", _closest_points_fp32)):
        with monkeypatch.context() as m:
            m.setattr(mod, attr, repl)
            _, res = _solve()
        # the 7-wire counts survive ...
        assert [(c.n_attempts, c.n_active, c.converged) for c in res.calls] == \
            [(c.n_attempts, c.n_active, c.converged) for c in ref.calls], name
        devs[name] = _max_dev(res, ref)
    print("largest relative change of ||R|| / f_ref per attempt:", devs)
    # ... but the residual moves by orders of magnitude more than 1e-6
    assert devs["element force"] > 1e-5 and devs["pair geometry"] > 1e-5, devs
