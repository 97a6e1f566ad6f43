"""ctypes binding of ``libxkamd.so`` (declarations in include/xkamd.h).

The library is built in-tree by ``build.py``.  There is no fallback: if the
shared object is missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxkamd.so")

_lib = None

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)
c_int64_p = C.POINTER(C.c_int64)
c_int8_p = C.POINTER(C.c_int8)
vp = C.c_void_p

PAIR_GEOM, PAIR_HYST, PAIR_HUBER, PAIR_FRIC, PAIR_TAN_C, PAIR_TAN_F = 1, 2, 4, 8, 16, 32
PAIR_COAT_F, PAIR_COAT_FRIC, PAIR_COAT_K, PAIR_COAT_KF = 128, 256, 512, 1024
PREC = {"fp64": 0, "mixed": 1, "fp32": 2}
OPT_STAGED = 256
OPT_HOST_PATTERN = 512
FLAG_COAT_UNPLACED = 8
# device Newton control record (XK_NC_* of include/xkamd.h)
NC_STATE, NC_ATT, NC_REASON, NC_NACT, NC_RATIO, NC_DU, NC_U, NC_ENERGY, NC_REL, NC_ALPHA = range(10)
NC_HIST, NC_HSTRIDE = 32, 4
NC_RUNNING, NC_CONVERGED, NC_DIVERGED, NC_STOPPED, NC_ASSIST, NC_REBUILD = range(6)
NCR_NONFINITE, NCR_DEGENERATE, NCR_SINGULAR, NCR_SOLVE, NCR_LIMIT = 1, 2, 4, 8, 16


class XkTuning(C.Structure):
    _fields_ = [("cell_factor", C.c_double), ("pair_block", C.c_int), ("gather_block", C.c_int)]


class XkDyn(C.Structure):
    _fields_ = [("M_diag", vp), ("u_pred", vp), ("v_pred", vp), ("acc_old", vp), ("acc", vp), ("vel", vp),
                ("dt", C.c_double), ("alpha_m", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


# name: (restype, argtypes)
SIGNATURES = {
    "xk_ctx_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, vp, vp, vp, C.c_double, vp, vp, C.c_int]),
    "xk_ctx_destroy": (C.c_int, [vp]),
    "xk_last_error": (C.c_char_p, []),
    "xk_ctx_precision": (C.c_int, [vp]),
    "xk_ctx_direct": (C.c_int, [vp]),
    "xk_beam_assemble": (C.c_int, [vp, vp, vp, vp, C.c_int64, vp]),
    "xk_pairs_count": (C.c_int, [vp, c_int64_p]),
    "xk_pairs_set": (C.c_int, [vp, C.c_int64] + [vp] * 15),
    "xk_pairs_get": (C.c_int, [vp] + [vp] * 15),
    "xk_detect_candidates": (C.c_int, [vp, vp, C.c_double, C.c_double, vp, C.c_double, c_int64_p, c_int64_p, vp]),
    "xk_broadphase": (C.c_int, [vp, vp, C.c_double, C.c_double, vp, C.c_int64, c_int64_p, vp]),
    "xk_last_candidates": (C.c_int, [vp, vp, C.c_int64]),
    "xk_strategy_detect": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_int, c_int64_p, vp]),
    "xk_pairs_constraint_jacobian": (C.c_int, [vp, C.c_int, C.c_int64, c_int64_p, c_int64_p, vp, vp, vp, vp,
                                               C.c_int64, C.c_int64, vp]),
    "xk_prescreen_radii": (C.c_int, [vp, vp, C.c_double, C.c_double, vp, vp, C.c_int64, c_int64_p, vp]),
    "xk_pairs_run": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_int, vp]),
    "xk_pairs_gather": (C.c_int, [vp, vp, vp, C.c_int64, vp]),
    "xk_pattern_build": (C.c_int, [vp]),
    "xk_pattern_info": (C.c_int, [vp, c_int64_p, c_int64_p, c_int64_p]),
    "xk_pattern_export": (C.c_int, [vp, vp, vp]),
    "xk_assemble_iteration": (C.c_int, [vp, vp, vp, vp, vp, C.POINTER(XkDyn), C.c_double, C.c_double,
                                        C.c_double, C.c_double, vp, vp, C.c_int64, vp, vp, vp]),
    "xk_linear_solve": (C.c_int, [vp, vp, vp, C.POINTER(XkDyn), vp, vp, C.c_int, C.c_double, vp,
                                  C.POINTER(C.c_int32), vp]),
    "xk_solve_buffer": (C.c_int, [vp, C.c_int, vp, C.c_int64, c_int64_p]),
    "xk_solve_set_lu": (C.c_int, [vp, C.c_int, c_int32_p, c_int64_p]),
    "xk_newton_update": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_double, C.c_double, vp, vp]),
    "xk_newton_control_begin": (C.c_int, [vp, vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_double, vp]),
    "xk_newton_control_attach": (C.c_int, [vp, vp]),
    "xk_newton_control_force": (C.c_int, [vp, vp, vp, vp]),
    "xk_newton_control_set": (C.c_int, [vp, C.c_int, C.c_int, vp]),
    "xk_status": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "xk_profile_begin": (C.c_int, [vp, C.c_int]),
    "xk_profile_end": (C.c_int, [vp, vp, C.POINTER(C.c_int)]),
    "xk_ctx_set_friction": (C.c_int, [vp, vp]),
    "xk_ctx_set_tuning": (C.c_int, [vp, C.POINTER(XkTuning)]),
    "xk_ctx_get_tuning": (C.c_int, [vp, C.POINTER(XkTuning)]),
    "xk_lp_set_params": (C.c_int, [vp, vp, C.c_int, C.c_float]),
    "xk_lp_features": (C.c_int, [vp, vp, vp, vp, C.c_float, vp, vp]),
    "xk_lp_select": (C.c_int, [vp, c_int64_p, vp]),
    "xk_lp_verify": (C.c_int, [vp, vp, vp, C.c_double, c_int64_p, vp]),
    "xk_lp_off": (C.c_int, [vp]),
    "xk_ctx_set_coating": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, vp]),
    "xk_coating_set_dt": (C.c_int, [vp, C.c_double]),
    "xk_coating_probe": (C.c_int, [vp, vp, c_int64_p, vp]),
    "xk_coating_probe_control": (C.c_int, [vp, vp, vp]),
    "xk_coating_commit": (C.c_int, [vp, vp]),
    "xk_pairs_get_coating": (C.c_int, [vp] + [vp] * 8),
    "xk_pairs_set_coating": (C.c_int, [vp] + [vp] * 6),
    "xk_pairs_snapshot": (C.c_int, [vp]),
    "xk_pairs_restore": (C.c_int, [vp]),
    "xk_mlp_n_params": (C.c_int, []),
    "xk_mlp_grad_work": (C.c_int64, [C.c_int64]),
    "xk_mlp_forward": (C.c_int, [vp, C.c_int64, vp, vp, vp]),
    "xk_mlp_grad": (C.c_int, [vp, vp, C.c_int64, vp, C.c_float, vp, vp, C.c_int64, vp]),
    "xk_mlp_adam": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, vp]),
    "xk_mlp_grad_noalloc": (C.c_int, [vp, vp, C.c_int64, vp, C.c_float, vp, vp]),
}


def lib():
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        # torch first: its bundled HIP runtime (soname libamdhip64.so.7) is then the
        # one this library binds to, so streams and device pointers are shared
        import torch  # noqa: F401

        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = lib().xk_last_error()
        raise RuntimeError(f"{what or 'xkamd'} failed: {msg.decode() if msg else 'unknown error'}")
    return rc


def ptr(a) -> int | None:
    """Raw address of a torch tensor / numpy array (None for None)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
