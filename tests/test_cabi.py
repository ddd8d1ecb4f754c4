"""The C-ABI library loads on the host and exports every entry point of include/portal_hip.h.

No compute calls are made here (no GPU in this environment); the host-side species
tables the library builds (Wigner symbols and L&L04 7.20 constants) are checked
against exact values through the library's debug exports.
"""
import ctypes as C
import json
import os
import re

import pytest

from portal_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "portal_hip.h")).read()
    return sorted(set(re.findall(r"\b(portal_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(L.EXPORTS) == names


def test_ctx_create_without_gpu_fails_cleanly():
    import torch  # noqa: F401  (only to know whether a device is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = L.load()
    h = C.c_void_p()
    assert lib.portal_ctx_create(C.byref(h), 0) == -2  # PORTAL_ERR_DEVICE


def test_library_wigner_and_coupling_tables(golden_dir):
    lib = L.load()
    lib.portal_debug_wig3jj.restype = C.c_double
    lib.portal_debug_wig6jj.restype = C.c_double
    d = json.load(open(os.path.join(golden_dir, "wigner_exact.json")))
    for j1, j2, j3, m1, m2, m3, v in d["w3j"]:
        got = lib.portal_debug_wig3jj(*[2 * int(x) for x in (j1, j2, j3, m1, m2, m3)])
        assert got == pytest.approx(v, rel=1e-14, abs=1e-15)
    for a, b, c, e, f, g, v in d["w6j"]:
        got = lib.portal_debug_wig6jj(*[2 * int(x) for x in (a, b, c, e, f, g)])
        assert got == pytest.approx(v, rel=1e-14, abs=1e-15)
    c0, c2 = C.c_double(), C.c_double()
    for kind, j, k, j2, k2, v0, v2 in d["coupling"]:
        lib.portal_debug_coupling(kind, j, k, j2, k2, C.byref(c0), C.byref(c2))
        assert c0.value == pytest.approx(v0, rel=1e-13, abs=1e-14)
        assert c2.value == pytest.approx(v2, rel=1e-13, abs=1e-14)
    for j, k, ju, ku, v in d["t_spont"]:
        lib.portal_debug_coupling(4, j, k, ju, ku, C.byref(c0), C.byref(c2))
        assert c0.value == pytest.approx(v, rel=1e-13, abs=1e-14)
