"""The C-ABI library: loads without a GPU and exports every declared entry point.

No compute call is made here (that needs a GPU); the GPU tests exercise
every entry point through the same binding.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from asl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "aslhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(aslhip_\w+)\s*\(", text)))


def test_header_parses():
    names = declared()
    assert "aslhip_step" in names and "aslhip_lattice_create" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert sorted(_lib.SIGNATURES) == declared()


def test_error_path_without_gpu_or_bad_args():
    # bad arguments are rejected before any device work, with ASL's prefix
    lib = _lib.load_library()
    h = ctypes.c_void_p()
    r = lib.aslhip_lattice_create(ctypes.byref(h), 8, 8, 8, 9, 0, 0)
    assert r == -1
    assert _lib.last_error().startswith("ASL ERROR")
    r = lib.aslhip_lattice_create(ctypes.byref(h), 0, 8, 8, 19, 0, 0)
    assert r == -1
    r = lib.aslhip_lattice_create(ctypes.byref(h), 8, 8, 8, 19, 0, 3)
    assert r == -1 and "periodic" in _lib.last_error()
    with pytest.raises(_lib.ASLError):
        _lib.call("aslhip_step", None, 1, 0)


def test_oracle_is_not_linked_into_the_product():
    # the product library must not depend on the oracle
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"orc_" not in data and b"lbgk_oracle" not in data


def test_template_entry_point_matches_oracle_templates():
    """aslhip_template (a host-only call) reports the same tables the oracle is pinned on."""
    import asl_amd
    from oracle import template
    for q in (15, 19, 27):
        t = asl_amd.VectorTemplate(q)
        c, w, inv = template(q)
        assert np.array_equal(t.c, c) and np.array_equal(t.inv, inv)
        np.testing.assert_allclose(t.w, w, rtol=0, atol=1e-16)
