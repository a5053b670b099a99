"""aslhip_read_f_noalloc: the populations' host view through the library's
fixed staging array instead of a buffer allocated per call.  It must return
what aslhip_read_f returns, bit for bit, on consecutive calls (the staging
array keeps the previous call's contents), on a lattice whose view spans
several staging pieces and then on a smaller one.
"""
import ctypes

import numpy as np
import pytest

import asl_amd
from asl_amd import _lib

pytestmark = pytest.mark.gpu


def channel(size, dtype, steps):
    L = asl_amd.Lattice(size, 19, dtype=dtype, nu=0.05)
    L.init_equilibrium(1.0, (0.03, 0.01, 0.0))
    ids = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0.0, 0.0)),
           L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
           L.add_slice_bc(_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE)]
    for i in ids:
        L.execute_bc(i)
    L.step(steps)
    return L


def read_noalloc(L):
    out = np.empty(L.q * L.len, dtype=L.host_dtype)
    _lib.call("aslhip_read_f_noalloc", L.h, out.ctypes.data)
    return out.reshape((L.q,) + L.S)


def test_noalloc_read_equals_read_f():
    big = channel((70, 70, 52), "f32", 37)  # 19 x 72 x 72 x 54 x 4 B = 21 MB: two staging pieces
    assert big.q * big.len * 4 > _lib.NOALLOC_SCRATCH_BYTES
    want = big.f()
    got = read_noalloc(big)
    assert np.array_equal(got, want, equal_nan=True)
    small = channel((12, 10, 8), "f64", 5)  # the next call: a smaller view, another element type
    want = small.f()
    got = read_noalloc(small)
    assert np.array_equal(got, want, equal_nan=True)
    big.step(3, defer=True)  # a pending step runs first, as for read_f
    got = read_noalloc(big)
    assert np.array_equal(got, big.f(), equal_nan=True)
