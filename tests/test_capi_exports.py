"""The C ABI library loads on a machine without a GPU and exports every entry
point include/bfpsh.h declares; the Python binding binds exactly those.
No compute calls are made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "bfpsh.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bfpsh_\w+)\s*\(", src, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from bfps_amd import build, capi
    if not os.path.exists(capi.LIB):
        build.build(verbose=False)
    return ctypes.CDLL(capi.LIB)


def test_header_declares_entry_points():
    names = declared()
    assert "bfpsh_step" in names and "bfpsh_particles_complete_loop" in names
    assert len(names) >= 20


def test_library_exports_all_declared(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from bfps_amd import capi
    assert sorted(capi.SIGNATURES) == declared()


def test_binding_loads(lib):
    from bfps_amd import capi
    L = capi.load()
    for n in declared():
        assert getattr(L, n).argtypes is not None
