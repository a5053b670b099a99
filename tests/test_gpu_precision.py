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
