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


def test_create_rejects_bad_arguments_before_any_device_call():
    """askja_create validates N, D, extents and precision on the host and
    leaves *out NULL (no device call: runs without a GPU).  The reference
    reads these from the params file unchecked (sun.cpp:803-938)."""
    import ctypes

    import askja_amd
    lib = askja_amd.lib()
    EINVAL, EUNSUP = -1, -4

    def create(nc, dim, ext, prec=0):
        arr = (ctypes.c_int * max(1, len(ext)))(*ext)
        out = ctypes.c_void_p(1234)
        rc = lib.askja_create(nc, dim, arr, prec, ctypes.byref(out))
        return rc, out.value

    for args, want in (
        ((0, 2, [4, 4]), EUNSUP),           # N < 1
        ((65, 2, [4, 4]), EUNSUP),          # N > 64
        ((3, 0, [4]), EUNSUP),              # D < 1
        ((3, 33, [2] * 33), EUNSUP),        # D > 32
        ((3, 2, [4, 4], 7), EUNSUP),        # unknown precision
        ((3, 2, [4, 4], -1), EUNSUP),
        ((3, 3, [4, 0, 4]), EINVAL),        # an extent < 1
        ((3, 2, [-2, 4]), EINVAL),
        ((3, 4, [1024, 1024, 1024, 2]), EUNSUP),  # more than 2^30 links
    ):
        rc, out = create(*args)
        assert rc == want, (args, rc)
        assert out is None, args
    null_out = lib.askja_create(3, 2, (ctypes.c_int * 2)(4, 4), 0, None)
    assert null_out == EINVAL


def test_integration_guide_lists_every_entry_point():
    """INTEGRATION.md's entry-point index names every function the header
    declares, so the guide cannot fall behind the ABI."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    missing = [n for n in declared_functions() if f"`{n}`" not in text]
    assert not missing, missing
