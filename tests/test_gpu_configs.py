"""The five measured workloads at their full sizes, on the GPU.

* 64^3 lid-driven cavity, 1000 FP32 steps: FP32 against the oracle (bit for
  bit) and against an FP64 GPU run (error stated);
* 256x128x128 sphere channel with the seeded inflow perturbation: FP32
  against the oracle after 100 steps (bit for bit) and against FP64 after
  the full 10,000 steps (error stated);
* 512^3 D3Q27 periodic Taylor-Green vortex at tau = 0.5005: FP32, FP16 and
  BF16 storage against FP64 after 500 steps, ranked and bounded;
* 64 seeded lid-driven cavities (edge log-uniform in [32, 160], tau uniform
  in [0.51, 1]) in one batch with 100-step graph chunks: every member equal
  to its own single-lattice run, two members equal to the oracle;
* the learned closure on the 128^3 periodic decaying-turbulence lattice:
  matrix-core evaluation against the oracle's BF16 model.

Each test prints the figures it checks (pytest -s shows them).
"""
import numpy as np
import pytest

import asl_amd
import bench
import oracle
from asl_amd import _lib, closure

pytestmark = pytest.mark.gpu
WALLS = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0
INNER = (slice(1, -1),) * 3


def cavity(n, tau, dtype="f32", lid=0.05):
    L = asl_amd.Lattice((n, n, n), 19, dtype=dtype, nu=(tau - 0.5) / 3.0)
    L.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    for b in (L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(lid, 0.0, 0.0)),
              L.add_slice_bc(_lib.BC_NOSLIP, WALLS)):
        L.execute_bc(b)
    return L


def oracle_cavity(n, tau, fp64=False, lid=0.05):
    O = oracle.OracleLattice((n, n, n), 19, fp64=fp64, nu=(tau - 0.5) / 3.0)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, _lib.ZE, u=(lid, 0.0, 0.0)), O.add_slice_bc(O.NOSLIP, WALLS)]
    for b in ids:
        O.execute_bc(b)
    return O, ids


def rel_errors(G, R, mask):
    """max|d rho| and max|d u| / max|u_R| over mask (R: the FP64 or oracle fields)."""
    rg, rr = G.rho().astype(np.float64), R.rho().astype(np.float64)
    ug, ur = G.velocity().astype(np.float64), R.velocity().astype(np.float64)
    drho = float(np.abs(rg[mask] - rr[mask]).max())
    umax = float(max(np.abs(ur[d][mask]).max() for d in range(3)))
    du = float(max(np.abs(ug[d][mask] - ur[d][mask]).max() for d in range(3)))
    return drho, du / umax, umax


def test_cavity_64_fp32_vs_oracle_and_fp64():
    n, tau, steps = 64, 0.6, 1000
    G32, G64 = cavity(n, tau), cavity(n, tau, "f64")
    G32.step(steps, graph=True)
    G64.step(steps, graph=True)
    O, ids = oracle_cavity(n, tau)
    O.run(steps, ids)
    mask = np.zeros(G32.S, dtype=bool)
    mask[INNER] = True
    dr_o, du_o, _ = rel_errors(G32, O, mask)
    dr_64, du_64, umax = rel_errors(G32, G64, mask)
    print("64^3 cavity, 1000 FP32 steps: vs oracle max|drho| %.3e, max|du|/max|u| %.3e; "
          "vs FP64 GPU max|drho| %.3e, max|du|/max|u| %.3e (max|u| %.4f)" % (dr_o, du_o, dr_64, du_64, umax))
    assert dr_o == 0.0 and du_o == 0.0
    assert dr_64 < 1e-4 and du_64 < 1e-3


def sphere_pair(storage):
    mapv, u = bench.setup_fields()
    return mapv, u, bench.build_gpu(mapv, u, storage)


def test_sphere_channel_full_size():
    mapv, u, G = sphere_pair("f32")
    O = oracle.OracleLattice(bench.SIZE, 19, fp64=False, nu=bench.NU)
    O.init_f(1.0, u.astype(np.float64))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, _lib.X0, u=(bench.U_IN, 0.0, 0.0)),
           O.add_slice_bc(O.CONSTANT_PRESSURE, _lib.XE, p=1.0),
           O.add_slice_bc(O.NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
           O.add_map_bc(O.NOSLIP_MAP, mapv)]
    for b in ids:
        O.execute_bc(b)
    G.step(100, graph=True)
    O.run(100, ids)
    ct = G.cell_types()
    mask = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    dr_o, du_o, _ = rel_errors(G, O, mask)
    print("sphere channel 256x128x128, 100 steps with the seeded perturbation: vs oracle max|drho| %.3e, "
          "max|du|/max|u| %.3e over %d computation nodes" % (dr_o, du_o, int(mask.sum())))
    assert dr_o == 0.0 and du_o == 0.0
    del O
    _, _, G64 = sphere_pair("f64")
    G.step(9900, graph=True)
    G64.step(10000, graph=True)
    dr, du, umax = rel_errors(G, G64, mask)
    print("sphere channel, 10,000 steps: FP32 vs FP64 GPU max|drho| %.3e, max|du|/max|u| %.3e (max|u| %.4f)"
          % (dr, du, umax))
    assert np.isfinite(dr) and dr < 1e-3 and du < 1e-2


def tgv_run(n, steps, dtype, nu, u0):
    rho, u = bench.tgv_fields(n, u0)
    cast = np.float64 if dtype == "f64" else np.float32
    L = asl_amd.Lattice((n, n, n), 27, dtype=dtype, periodic=True, nu=nu)
    L.init_equilibrium(rho.astype(cast, copy=False), u.astype(cast, copy=False))
    del rho, u
    ke0 = L.conservation()[4]
    L.step(steps, graph=True)
    ke = L.conservation()[4]
    v = L.velocity()
    L.close()
    return v, ke0, ke


def test_precision_tgv_512_d3q27():
    n, steps, u0 = 512, 500, 0.04
    nu = (0.5005 - 0.5) / 3.0
    ref, ke0_64, ke64 = tgv_run(n, steps, "f64", nu, u0)
    err = {}
    for dt in ("f32", "f16", "bf16"):
        v, ke0, ke = tgv_run(n, steps, dt, nu, u0)
        du = 0.0
        for d in range(3):
            du = max(du, float(np.abs(v[d][INNER].astype(np.float64) - ref[d][INNER]).max()))
        del v
        err[dt] = (du / u0, abs(ke / ke0 - ke64 / ke0_64))
        print("D3Q27 TGV 512^3, tau 0.5005, %d steps: %s storage vs FP64 max|du|/u0 %.3e, |d(KE/KE0)| %.3e"
              % (steps, dt, err[dt][0], err[dt][1]))
    assert err["f32"][0] < err["f16"][0] < err["bf16"][0]
    assert err["f32"][0] < 1e-3 and err["f16"][0] < 1e-2 and err["bf16"][0] < 1e-1
    assert err["f32"][1] < 1e-4 and err["f16"][1] < 1e-3 and err["bf16"][1] < 3e-2


def test_batch_of_64_cavities():
    specs = bench.batch_specs()
    assert len(specs) == 64 and min(n for n, _ in specs) >= 32 and max(n for n, _ in specs) <= 160
    steps = 250  # two graph chunks and a 50-step remainder
    members = [cavity(n, t) for n, t in specs]
    B = asl_amd.Batch(members)
    B.step(steps, graph=True)
    diffs = 0
    for (n, t), M in zip(specs, members):
        S = cavity(n, t)
        S.step(steps, graph=True)
        same = (np.array_equal(M.f()[(slice(None),) + INNER], S.f()[(slice(None),) + INNER])
                and np.array_equal(M.rho()[INNER], S.rho()[INNER])
                and np.array_equal(M.velocity()[(slice(None),) + INNER], S.velocity()[(slice(None),) + INNER]))
        diffs += 0 if same else 1
        S.close()
    print("64 cavities (%d cells), %d graph steps: %d members differ from their single-lattice runs"
          % (sum(n ** 3 for n, _ in specs), steps, diffs))
    assert diffs == 0
    order = sorted(range(64), key=lambda i: specs[i][0])
    for i in (order[0], order[32]):
        n, t = specs[i]
        O, ids = oracle_cavity(n, t)
        O.run(steps, ids)
        mask = np.zeros(members[i].S, dtype=bool)
        mask[INNER] = True
        dr, du, _ = rel_errors(members[i], O, mask)
        print("batch member %d (L=%d, tau=%.4f) vs oracle: max|drho| %.3e, max|du|/max|u| %.3e" % (i, n, t, dr, du))
        assert dr == 0.0 and du == 0.0


def test_closure_128_matrix_cores_vs_oracle():
    n = 128
    nu = (0.51 - 0.5) / 3.0 / 4.0  # the coarse viscosity of a 4x finer lattice at tau 0.51
    u = closure.with_ghosts(closure.turbulence_field(n, bench.SEED, 0.03))
    rho = closure.with_ghosts(np.float32(1.0) + np.float32(3.0) * closure.pressure_field(u[(slice(None),) + INNER]))
    G = asl_amd.Lattice((n, n, n), 19, dtype="f32", periodic=True, nu=nu)
    G.init_equilibrium(rho, u)
    O = oracle.OracleLattice((n, n, n), 19, fp64=False, nu=nu, periodic=True)
    O.init_f(rho.astype(np.float64), u.astype(np.float64))
    p = closure.init_params(5)
    p[-1] = 0.3
    args = (p, np.zeros(7, np.float32), np.full(7, 40.0, np.float32), 0.05)
    G.set_closure(*args)
    O.set_closure(*args)
    s = (slice(None),) + INNER
    G.step(1)
    O.run(1)
    d1 = float(np.abs(G.f()[s].astype(np.float64) - O.f()[s]).max())
    G.step(4)
    O.run(4)
    d5 = float(np.abs(G.f()[s].astype(np.float64) - O.f()[s]).max())
    print("closure on 128^3 decaying turbulence: BF16 matrix cores vs oracle max|df| after 1 step %.2e, "
          "after 5 steps %.2e" % (d1, d5))
    assert d1 < 2e-6 and d5 < 1e-5
