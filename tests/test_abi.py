"""The C-ABI library loads and exports every entry point include/askja.h
declares, with the Python binding's signature table covering all of them
(no device calls: runs without a GPU)."""
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "askja.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(askja_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("askja_create", "askja_kick", "askja_drift", "askja_action", "askja_hmc_trajectory",
                 "askja_wilson_lines", "askja_mom_gsl", "askja_reunitarize"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import askja_amd
    if not os.path.exists(askja_amd.LIB_PATH):
        pytest.fail("libaskja.so is not built (run __graft_entry__.build())")
    lib = askja_amd.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(askja_amd.SIGNATURES)


def test_product_path_has_no_fallback(monkeypatch):
    """Without the built library the binding raises instead of computing."""
    import askja_amd
    monkeypatch.setattr(askja_amd, "_LIB", None)
    monkeypatch.setattr(askja_amd, "LIB_PATH", "/nonexistent/libaskja.so")
    with pytest.raises(askja_amd.AskjaError):
        askja_amd.Lattice(3, [2, 2])
