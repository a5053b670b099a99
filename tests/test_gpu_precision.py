"""Reduced-precision storage of the populations (BF16 / FP16 of f - 1).

D3Q27 periodic Taylor-Green vortex: FP32, FP16 and BF16 storage against an
FP64 run of the same lattice.  ASL has no reduced-precision storage, so
there is no ASL result to match; the FP64 run is the reference (its own
parity with the oracle is tested elsewhere) and the error must rank with
the storage precision: FP32 < FP16 < BF16, each within its bound.
"""
import numpy as np
import pytest

import asl_amd
import bench

pytestmark = pytest.mark.gpu


def run(n, steps, dtype, nu, u0):
    rho, u = bench.tgv_fields(n, u0)
    L = asl_amd.Lattice((n, n, n), 27, dtype=dtype, periodic=True, nu=nu)
    L.init_equilibrium(rho, u)
    ke0 = L.conservation()[4]
    L.step(steps)
    return L.velocity()[(slice(None),) + (slice(1, -1),) * 3].astype(np.float64), ke0, L.conservation()[4]


def test_storage_precision_ranks_against_fp64():
    n, steps, u0 = 48, 300, 0.04
    nu = (0.51 - 0.5) / 3.0
    ref, ke0_64, ke64 = run(n, steps, "f64", nu, u0)
    err = {}
    for dt in ("f32", "f16", "bf16"):
        u, ke0, ke = run(n, steps, dt, nu, u0)
        err[dt] = (np.abs(u - ref).max() / u0, abs(ke / ke0 - ke64 / ke0_64))
        print("D3Q27 TGV %d^3 x %d, %s storage vs FP64: max|du|/u0 %.3e, |dKE/KE0| %.3e"
              % (n, steps, dt, err[dt][0], err[dt][1]))
    # measured on MI355X: 1.6e-5 (FP32), 7.4e-4 (FP16), 5.6e-3 (BF16); KE 1.5e-5, 2.4e-5, 8.1e-4
    assert err["f32"][0] < 1e-4
    assert err["f32"][0] < err["f16"][0] < err["bf16"][0]
    assert err["f16"][0] < 2e-3 and err["bf16"][0] < 2e-2
    assert err["f16"][1] < 1e-3 and err["bf16"][1] < 1e-2


def bf16_rne(x):
    """Round float32 values to BF16 (nearest, ties to even), as float32 (numpy)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_storage_rounding_is_nearest_even(dtype):
    """write_f stores f - 1 rounded to nearest, ties to even (FP16: the IEEE
    conversion; BF16: the hardware conversion), read_f returns 1 + that."""
    size = (10, 9, 8)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(4)
    f = (1.0 + rng.standard_normal((19,) + S) * 10.0 ** rng.uniform(-6, -1, (19,) + S)).astype(np.float32)
    if dtype == "bf16":
        # exact BF16 ties in f - 1 (a BF16 value plus half its last place), where 1 + tie is exact
        dev = (f[:6] - np.float32(1.0)).astype(np.float32)
        tie = (bf16_rne(dev).view(np.uint32) | np.uint32(0x8000)).view(np.float32)
        ok = (np.abs(tie) > 1e-4) & ((np.float32(1.0) + tie) - np.float32(1.0) == tie)
        f[:6] = np.where(ok, np.float32(1.0) + tie, f[:6])
        assert ok.sum() > 1000
    L = asl_amd.Lattice(size, 19, dtype=dtype, nu=0.05)
    L.write_f(f)
    dev = (f - np.float32(1.0)).astype(np.float32)  # exact: f lies in [1/2, 2]
    stored = bf16_rne(dev) if dtype == "bf16" else dev.astype(np.float16).astype(np.float32)
    assert np.array_equal(L.f(), np.float32(1.0) + stored)
