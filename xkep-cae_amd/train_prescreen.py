"""Train the learned prescreen on the method's own runs and store the weights.

Runs dynamic radial squeezes (``workloads.squeeze_case``) of 7- and 19-wire
strands through the Newton driver on the device, full path, logging every
Newton call's increment-start features and the exact end-of-call gaps; trains
the network with the library's kernels; sets the keep threshold below the
lowest positive logit; writes ``prescreen_weights.npz`` next to this file.

Usage (on a GPU host):  python -m xkep-cae_amd.train_prescreen   (or run the file)
"""

from __future__ import annotations

import importlib
import os
import sys

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    __package__ = "xkep-cae_amd"

P = importlib.import_module(__package__ + ".prescreen")
S = importlib.import_module(__package__ + ".solver")
C = importlib.import_module(__package__ + ".context")
E = importlib.import_module(__package__ + ".elements")
K = importlib.import_module(__package__ + ".contact")
W = importlib.import_module(__package__ + ".workloads")

TRAIN_RUNS = [(f, nw) for f in (60.0, 80.0, 100.0, 120.0, 140.0) for nw in (7, 19)]


def run_case(case, *, precision: str = "fp64", net=None, log: bool = False, device_solve: bool = False,
             use_graph: bool = False, device_control: bool = True, coating: dict | None = None, core_radii=None):
    """Solve one squeeze case on the device; returns (SolverResult, backend).
    device_solve / use_graph: the Newton attempts on the device (linear solve,
    line search, update), after each increment's first as HIP-graph replays;
    device_control: their convergence and divergence tests on the device too;
    coating: ContactConfig's coating fields, with the pairs' core radii."""
    r = float(case["radii"][0])
    sec = C.circle_section(r, W.E_STEEL, W.NU_STEEL)
    ctx = C.DeviceContext(case["X"], case["conn"], sec, radii=case["radii"], precision=precision)
    Md = E.ULCRBeamAssembler(case["X"], case["conn"], sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y,
                             sec.kappa_z).lumped_mass_diag(W.RHO_STEEL)
    kcfg = K.ContactConfig(mu=case["mu"], k_t_ratio=case["ktr"], exclude_same_layer=True,
                           elem_layer_map=W.layer_map(case["layer"]), **(coating or {}))
    b = S.DeviceBackend(ctx, M_diag=Md, rho_inf=case["rho_inf"], k_pen=case["k_pen"], smoothing_delta=case["sd"],
                        mu=case["mu"], k_t_ratio=case["ktr"], fixed_dofs=case["fixed"], config=kcfg,
                        device_solve=device_solve, use_graph=use_graph, device_control=device_control,
                        core_radii=core_radii)
    if net is not None:
        b.use_prescreen(net, elem_layer=case["layer"])
    if log:
        b.log_features(True, elem_layer=case["layer"])
    res = S.ContactFrictionSolver(b, fixed_dofs=case["fixed"], f_ext_total=case["f"], t_total=case["t_total"],
                                  config=S.SolverConfig(**case["cfg"])).solve()
    return res, b


def dataset(runs=TRAIN_RUNS):
    Xs, ys = [], []
    for force, nw in runs:
        _, b = run_case(W.squeeze_case(force, n_wires=nw), log=True)
        for feat, gap in b.feature_log:
            x, y = P.labels_of(feat, gap)
            Xs.append(x)
            ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def train(X, y, *, epochs: int = 600, lr: float = 3e-3, seed: int = 0, margin: float = 1.0):
    net = P.PrescreenNet(seed=seed)
    hist = net.fit(X, y, epochs=epochs, lr=lr)
    net.choose_threshold(X, y, margin=margin)
    return net, hist


def main(out: str = P.WEIGHTS):
    X, y = dataset()
    net, hist = train(X, y)
    z = net.logits(X).cpu().numpy()
    kept = float(np.mean(z >= net.threshold))
    print(f"train rows {len(y)}, positives {int(y.sum())}, loss {hist[0]:.4f} -> {hist[-1]:.4f}, "
          f"threshold {net.threshold:.3f}, kept {kept:.3f} of non-contributing rows", flush=True)
    net.save(out)
    print(f"wrote {out}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else P.WEIGHTS)
