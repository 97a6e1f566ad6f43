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
