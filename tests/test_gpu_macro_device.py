"""Device addresses of rho and u for a solver coupled to the lattice on the
GPU (aslhip_macro_device): the memory holds exactly what the host reads
return, after eager, graph-replayed and deferred steps, for every storage
type."""
import ctypes

import numpy as np
import pytest

import asl_amd
from asl_amd import _lib

pytestmark = pytest.mark.gpu
HIP_MEMCPY_DEVICE_TO_HOST = 2


def _d2h(hip, ptr, n, dtype):
    out = np.empty(n, dtype=dtype)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                       HIP_MEMCPY_DEVICE_TO_HOST)
    assert rc == 0
    return out


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f64"])
def test_macro_device_holds_the_host_read(dtype):
    hip = ctypes.CDLL("libamdhip64.so")
    L = asl_amd.Lattice((20, 17, 22), 19, dtype=dtype, nu=0.1 / 3)
    L.init_equilibrium(1.0, (0, 0, 0))
    bcs = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)),
           L.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0)]
    for b in bcs:
        L.execute_bc(b)
    s = (slice(1, -1),) * 3
    for n, graph in ((7, False), (230, True)):
        L.step(n, graph=graph)
        pr, pu = L.macro_device()
        assert pr and pu
        rho = _d2h(hip, pr, L.len, L.host_dtype).reshape(L.S)
        u = _d2h(hip, pu, 3 * L.len, L.host_dtype).reshape((3,) + L.S)
        assert np.array_equal(rho[s], L.rho()[s])
        assert np.array_equal(u[(slice(None),) + s], L.velocity()[(slice(None),) + s])
        assert np.abs(u[(slice(None),) + s]).max() > 1e-3


def test_macro_device_after_deferred_execute():
    # asl::LBGK::execute() keeps its last step pending; the device fields must
    # show it, as a host read does
    hip = ctypes.CDLL("libamdhip64.so")
    blk = asl_amd.Block((16, 18, 14))
    lb = asl_amd.LBGK(blk, 0.05)
    lb.init()
    L = lb.lattice
    L.init_equilibrium(1.0, (0.02, 0, 0))
    bc = L.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE)
    L.execute_bc(bc)
    for _ in range(5):
        lb.execute()
        L.execute_bc(bc)
    pr, pu = L.macro_device()
    rho = _d2h(hip, pr, L.len, L.host_dtype).reshape(L.S)
    u = _d2h(hip, pu, 3 * L.len, L.host_dtype).reshape((3,) + L.S)
    s = (slice(1, -1),) * 3
    assert np.array_equal(rho[s], L.rho()[s])
    assert np.array_equal(u[(slice(None),) + s], L.velocity()[(slice(None),) + s])
    assert np.abs(u[0][s]).max() > 1e-3
