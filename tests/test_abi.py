"""The C-ABI library loads and exports every entry point include/nbpg.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared():
    src = open(os.path.join(ROOT, "include", "nbpg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nbpg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["nbpg_search", "nbpg_robust_prune_batch", "nbpg_compgt", "nbpg_index_create",
              "nbpg_l2_batch", "nbpg_build"]:
        assert n in names


def test_library_exports_all_symbols():
    path = os.path.join(ROOT, "nbpg_amd", "libnbpg_amd.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_binds_every_symbol():
    import nbpg_amd
    assert set(declared()) <= set(nbpg_amd._SIGS)


def test_product_does_not_import_oracle():
    """The shipped package never reaches the checker."""
    pkg = os.path.join(ROOT, "nbpg_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "nbpg_oracle" not in text, f
