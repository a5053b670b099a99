"""Device addresses of rho and u for a solver coupled to the lattice on the
GPU (aslhip_macro_device).

The consumer names its own stream: the library orders that stream after the
step that wrote the fields (no host wait) and orders the lattice's next step
after what the consumer queued before it.  The tests read the fields from the
device on a consumer stream and compare them with the CPU oracle's rho and u
after the same steps (eager, graph-replayed and deferred), and check that a
step requested right after the consumer queued its copy does not overwrite
the fields under it.
"""
import ctypes

import numpy as np
import pytest

import asl_amd
from asl_amd import _lib
from gpu_common import pair

pytestmark = pytest.mark.gpu
D2H, D2D = 2, 3
WALLS = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0


class Hip:
    """The few HIP runtime calls a consumer needs, through ctypes."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")

    def ok(self, rc):
        assert rc == 0, "HIP call failed (%d)" % rc

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.lib.hipStreamCreate(ctypes.byref(s)))
        return s.value

    def malloc(self, n):
        p = ctypes.c_void_p()
        self.ok(self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)))
        return p.value

    def free(self, p):
        self.ok(self.lib.hipFree(ctypes.c_void_p(p)))

    def copy_async(self, dst, src, n, kind, stream):
        self.ok(self.lib.hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), ctypes.c_size_t(n), kind,
                                        ctypes.c_void_p(stream)))

    def sync(self, stream):
        self.ok(self.lib.hipStreamSynchronize(ctypes.c_void_p(stream)))

    def read(self, ptr, n, dtype, stream):
        out = np.empty(n, dtype=dtype)
        self.copy_async(out.ctypes.data, ptr, out.nbytes, D2H, stream)
        self.sync(stream)
        return out


def cavity(dtype="f32"):
    G, O = pair((20, 17, 22), 19, nu=0.1 / 3, dtype=dtype)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    g = [G.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)), G.add_slice_bc(_lib.BC_NOSLIP, WALLS)]
    o = [O.add_slice_bc(O.CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)), O.add_slice_bc(O.NOSLIP, WALLS)]
    for a, b in zip(g, o):
        G.execute_bc(a)
        O.execute_bc(b)
    return G, O, o


def fields_on(hip, L, stream):
    pr, pu = L.macro_device(stream)
    assert pr and pu
    rho = hip.read(pr, L.len, L.host_dtype, stream).reshape(L.S)
    u = hip.read(pu, 3 * L.len, L.host_dtype, stream).reshape((3,) + L.S)
    return rho, u


def check_vs_oracle(rho, u, O, dtype):
    s = (slice(1, -1),) * 3
    ro, uo = O.rho()[s], O.velocity()[(slice(None),) + s]
    assert np.abs(uo).max() > 1e-3
    if dtype in ("f32", "f64"):  # the same operations in the same order: bit for bit
        assert np.array_equal(rho[s], ro.astype(rho.dtype))
        assert np.array_equal(u[(slice(None),) + s], uo.astype(u.dtype))
    else:  # 16-bit storage of f - 1 against the FP32 oracle
        assert np.abs(rho[s] - ro).max() < 5e-3
        assert np.abs(u[(slice(None),) + s] - uo).max() < 5e-3


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f64"])
def test_macro_device_equals_the_oracle(dtype):
    hip = Hip()
    st = hip.stream()
    G, O, o = cavity(dtype)
    for n, graph in ((7, False), (230, True)):
        G.step(n, graph=graph)
        O.run(n, o)
        rho, u = fields_on(hip, G, st)
        check_vs_oracle(rho, u, O, dtype)
        # and what the host reads
        s = (slice(1, -1),) * 3
        assert np.array_equal(rho[s], G.rho()[s])
        assert np.array_equal(u[(slice(None),) + s], G.velocity()[(slice(None),) + s])


def test_macro_device_after_deferred_execute():
    # asl::LBGK::execute() keeps its last step pending; the device fields show
    # it, as a host read does, and equal the oracle after the same 5 steps
    hip = Hip()
    st = hip.stream()
    G, O, o = cavity()
    for _ in range(5):
        G.step(1, defer=True)
        for b in (0, 1):
            G.execute_bc(b)
    O.run(5, o)
    rho, u = fields_on(hip, G, st)
    check_vs_oracle(rho, u, O, "f32")


def test_next_step_waits_for_the_consumer():
    """The consumer queues copies of rho and u on its stream; a step requested
    right after must not overwrite the fields before those copies ran."""
    hip = Hip()
    st = hip.stream()
    G, O, o = cavity()
    G.step(40)
    O.run(40, o)
    pr, pu = G.macro_device(st)
    nb = 3 * G.len * 4
    dst = [hip.malloc(nb) for _ in range(8)]
    try:
        for d in dst:  # several copies, so the consumer stream is busy for a while
            hip.copy_async(d, pu, nb, D2D, st)
        G.step(25, graph=False)  # writes rho and u
        hip.sync(st)
        s = (slice(None),) + (slice(1, -1),) * 3
        uo = O.velocity()[s].astype(np.float32)
        for d in dst:
            u = hip.read(d, 3 * G.len, np.float32, st).reshape((3,) + G.S)
            assert np.array_equal(u[s], uo)
    finally:
        for d in dst:
            hip.free(d)
    O.run(25, o)
    s = (slice(1, -1),) * 3
    assert np.array_equal(G.rho()[s], O.rho()[s].astype(np.float32))
