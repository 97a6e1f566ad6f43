"""The C-ABI library loads and exports every entry point include/xkamd.h
declares (no GPU calls)."""

import ctypes
import os
import re

from conftest import REPO, pkg


def declared():
    txt = open(os.path.join(REPO, "include", "xkamd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(xk_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path_entry_points():
    names = declared()
    for must in ["xk_ctx_create", "xk_beam_assemble", "xk_broadphase", "xk_detect_candidates", "xk_pairs_run",
                 "xk_pairs_gather", "xk_assemble_iteration", "xk_pattern_export"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = pkg("_lib")
    assert os.path.exists(L.LIB_PATH), "build the library first (__graft_entry__.build())"
    so = ctypes.CDLL(L.LIB_PATH)
    for name in declared():
        assert hasattr(so, name), name
    assert set(declared()) == set(L.SIGNATURES), "ctypes signatures out of sync with the header"
    L.lib()  # binds every signature


def test_product_package_does_not_import_the_oracle():
    root = os.path.join(REPO, "xkep-cae_amd")
    for dirpath, _, files in os.walk(root):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", src).replace("oracle/", ""), f


def _header_enum(prefix):
    """name -> value of the XK_<prefix>* enumerators of include/xkamd.h."""
    txt = open(os.path.join(REPO, "include", "xkamd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"\bXK_(" + prefix + r"[A-Z0-9_]*)\s*=\s*(\d+)", txt)}


def test_python_mirror_of_the_control_record_matches_the_header():
    """The device Newton control record's layout, states and reasons as the
    Python driver reads them (_lib.NC_*) are the header's XK_NC_* / XK_NCR_*."""
    L = pkg("_lib")
    nc = _header_enum("NC_")
    for name in ("STATE", "ATT", "REASON", "NACT", "RATIO", "DU", "U", "ENERGY", "REL", "ALPHA", "HIST", "HSTRIDE",
                 "RUNNING", "CONVERGED", "DIVERGED", "STOPPED", "ASSIST", "REBUILD"):
        assert getattr(L, "NC_" + name) == nc["NC_" + name], name
    ncr = _header_enum("NCR_")
    for name in ("NONFINITE", "DEGENERATE", "SINGULAR", "SOLVE", "LIMIT"):
        assert getattr(L, "NCR_" + name) == ncr["NCR_" + name], name
    # the record's scalar fields fit before the history
    assert max(v for k, v in nc.items() if k not in ("NC_HIST", "NC_HSTRIDE") and not k.startswith(
        ("NC_RUNNING", "NC_CONVERGED", "NC_DIVERGED", "NC_STOPPED", "NC_ASSIST", "NC_REBUILD"))) < nc["NC_HIST"]
    flags = _header_enum("FLAG_")
    S = pkg("solver")
    assert (S.FLAG_NONFINITE, S.FLAG_DEGENERATE, S.FLAG_SINGULAR) == (
        flags["FLAG_NONFINITE"], flags["FLAG_DEGENERATE"], flags["FLAG_SINGULAR"])
