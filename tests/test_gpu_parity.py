"""GPU parity: the HIP path (through the C ABI) vs the CPU oracle.

Both evaluate the same operations in the same order with the same
contractions, so results are compared bit for bit on computation nodes;
the contract of the port is ρ and u within 1e-5 (asserted as well).
FP32 results are also compared with FP64 runs (oracle and GPU) to state
the rounding error of the method itself.
"""
import numpy as np
import pytest

import asl_amd
from asl_amd import _lib, geometry
from gpu_common import add_bc, compare, pair, run

pytestmark = pytest.mark.gpu
ALL_BUT_ZE = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0


def cavity(n, steps, q=19, dtype="f32", lid_first=True, pre_execute=True, nu=(0.6 - 0.5) / 3):
    G, O = pair((n, n, n), q, nu=nu, dtype=dtype)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    lid = add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0))
    walls = add_bc(G, O, _lib.BC_NOSLIP, ALL_BUT_ZE)
    run(G, O, steps, [lid, walls] if lid_first else [walls, lid], pre_execute)
    return G, O


def test_cavity_64_1000_steps_fp32_bitwise():
    """The 64^3 lid-driven cavity (tau = 0.6, lid 0.05 on ZE), 1000 FP32 steps."""
    G, O = cavity(64, 1000)
    d = compare(G, O)
    print("cavity 64^3 x 1000 FP32 GPU vs oracle:", d)
    assert d["cells"] == 64 ** 3
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0
    assert d["rho"] <= 1e-5 and d["u"] <= 1e-5 * d["umax"]


def test_cavity_fp32_error_vs_fp64():
    """The FP32 rounding error of the method, against FP64 runs (oracle and GPU)."""
    G32, O32 = cavity(32, 400)
    G64, O64 = cavity(32, 400, dtype="f64")
    d64 = compare(G64, O64)
    assert d64["rho"] == 0.0 and d64["u"] == 0.0  # FP64 GPU == FP64 oracle
    m = np.zeros(G32.S, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    r32, r64 = G32.rho()[m].astype(float), G64.rho()[m]
    u32, u64 = G32.velocity()[:, m].astype(float), G64.velocity()[:, m]
    er = np.abs(r32 - r64).max()
    eu = np.abs(u32 - u64).max() / np.abs(u64).max()
    print("FP32 vs FP64 after 400 steps: max|drho| %.2e, max|du|/max|u| %.2e" % (er, eu))
    # measured on MI355X: 2.0e-5 and 2.3e-5 (the FP32 rounding of the method)
    assert er < 1e-4 and eu < 1e-4


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("lid_first", [True, False])
def test_cavity_templates_and_order(q, lid_first):
    G, O = cavity(12, 60, q=q, lid_first=lid_first)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


def test_first_step_reads_raw_ghosts_when_no_bc_was_executed():
    # ASL pulls the initial ghost values in the first step if no BC ran yet
    G, O = cavity(10, 25, pre_execute=False)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0, d


@pytest.mark.parametrize("deferred_steps", [1, 3])
def test_deferred_step_keeps_its_bc_order(deferred_steps):
    """asl::LBGK::execute() leaves its step pending; BCs executed after it in
    a changed order belong to the step after it.  Driver sequence: BCs (lid,
    walls); execute(); BCs (walls, lid); 30 more steps.  The pending step must
    still let the walls win the lid edges, as ASL's own order does."""
    G, O = pair((12, 11, 10), 19, nu=(0.6 - 0.5) / 3)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    lid = add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0))
    walls = add_bc(G, O, _lib.BC_NOSLIP, ALL_BUT_ZE)
    for g, o in (lid, walls):
        G.execute_bc(g)
        O.execute_bc(o)
    G.step(deferred_steps, defer=True)
    G.execute_bc(walls[0])
    G.execute_bc(lid[0])
    G.step(30)
    # the steps of one call repeat the active order between them
    O.run(deferred_steps - 1, [lid[1], walls[1]])
    O.run(1 + 30, [walls[1], lid[1]])
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


def test_pure_streaming_bitwise():
    # nu = inf: w = 0, populations only stream and bounce
    n = 12
    S = (n + 2,) * 3
    rng = np.random.default_rng(2)
    u = (0.05 * rng.standard_normal((3,) + S)).astype(np.float32)
    rho = (1 + 0.05 * rng.standard_normal(S)).astype(np.float32)
    G, O = pair((n, n, n), 19, nu=float("inf"))
    G.init_equilibrium(rho, u)
    O.init_f(rho.astype(float), u.astype(float))
    walls = add_bc(G, O, _lib.BC_NOSLIP, 63)
    run(G, O, 37, [walls])
    d = compare(G, O)
    assert d["f"] == 0.0 and d["rho"] == 0.0, d


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("compressible", [True, False])
def test_all_bc_kinds(q, compressible):
    size = (24, 14, 13)
    sphere = geometry.normalize(geometry.sphere(size, 3.2, (8.0, 6.5, 6.0)))
    pmap = geometry.normalize(geometry.sphere(size, 2.0, (17.0, 7.0, 6.0)))
    G, O = pair(size, q, nu=0.04, compressible=compressible)
    G.init_equilibrium(1.0, (0.03, 0, 0))
    O.init_f(1.0, (0.03, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0, 0)),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
           add_bc(G, O, _lib.BC_NOSLIP, _lib.Y0 | _lib.YE),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE_VELOCITY, _lib.Z0 | _lib.ZE, p=1.001, u=(0.02, 0.0, 0.0)),
           add_bc(G, O, _lib.BC_NOSLIP_MAP, mapv=sphere),
           add_bc(G, O, _lib.BC_PRESSURE_VELOCITY_MAP, mapv=pmap, p=1.002, u=(0.0, 0.01, 0.0))]
    run(G, O, 80, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


@pytest.mark.parametrize("rotating", [False, True, "block"])
def test_velocity_map(rotating):
    # BCVelocityMap: moving obstacle surface (constant or rotating wall); "block":
    # the rotation field on a lattice whose Block has dx 0.25 at (-3.1, 2.2, 0.7),
    # so the wall velocity is omega x (r - centre) at physical positions r
    size = (20, 16, 14)
    body = geometry.normalize(geometry.sphere(size, 3.5, (9.0, 7.5, 6.5)))
    G, O = pair(size, 19, nu=0.05)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    rot = (0.0, 0.0, 0.01, 9.0, 7.5, 6.5) if rotating else None
    if rotating == "block":
        dx, pos = 0.25, (-3.1, 2.2, 0.7)
        G.set_geometry(dx, pos)
        O.set_geometry(dx, pos)
        rot = (0.0, 0.01 / dx, 0.02 / dx, pos[0] + 8.0 * dx, pos[1] + 6.5 * dx, pos[2] + 5.5 * dx)
    bcs = [add_bc(G, O, _lib.BC_NOSLIP, 63),
           add_bc(G, O, _lib.BC_VELOCITY_MAP, mapv=body, u=(0.02, 0.0, 0.01), rot=rot)]
    run(G, O, 60, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d
    assert d["umax"] > 1e-3


def casing(size, radius):
    """The fluid inside a z-cylinder (a compressor casing): map = (radius - r)
    / 3, clipped to [-1, 1], on the ghost-inclusive grid (ASL layout); the
    guard band is three cells thick."""
    S = tuple(n + 2 for n in size)
    x = np.arange(S[0], dtype=np.float64)[:, None, None] - (S[0] - 1) / 2.0
    y = np.arange(S[1], dtype=np.float64)[None, :, None] - (S[1] - 1) / 2.0
    m = (radius - np.sqrt(x * x + y * y)) / 3.0 + np.zeros((1, 1, S[2]))
    return np.clip(m, -1.0, 1.0).astype(np.float32)


@pytest.mark.parametrize("q", [15, 19])
def test_velocity_map_skipping_outer_layer(q):
    # compressor.cc's set-up in small: a rotating velocity-map casing around
    # the fluid, pressure on Z0 / ZE.  The casing's boundary nodes in the first
    # and last z layer average the density over outer-layer neighbours, which
    # ASL reads from stale memory: refused by default, left out of the
    # average with the opt-in (the oracle restates the same rule)
    size = (16, 16, 12)
    m = casing(size, 4.6)
    rot = (0.0, 0.0, 0.01, 8.5, 8.5, 6.5)
    G, O = pair(size, q, nu=0.05)
    G.set_omega((0.0, 0.0, 0.002))
    O.set_omega((0.0, 0.0, 0.002))
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_VELOCITY_MAP, mapv=m, rot=rot),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE, _lib.Z0 | _lib.ZE, p=1.0)]
    with pytest.raises(asl_amd.ASLError):
        for g, _ in bcs:
            G.execute_bc(g)
        G.step(2)
    G, O = pair(size, q, nu=0.05)
    for L in (G, O):
        L.set_omega((0.0, 0.0, 0.002))
        L.set_vmap_skip_outer(True)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_VELOCITY_MAP, mapv=m, rot=rot),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE, _lib.Z0 | _lib.ZE, p=1.0)]
    run(G, O, 80, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d
    assert d["umax"] > 1e-3


@pytest.mark.parametrize("q", [19, 27])
def test_periodic_taylor_green(q):
    N, nu, u0 = 24, 0.02, 0.03
    size = (N, N, N)
    S = tuple(s + 2 for s in size)
    k = 2 * np.pi / N
    x = (np.arange(S[0]) - 1.0)[:, None, None]
    y = (np.arange(S[1]) - 1.0)[None, :, None]
    z = (np.arange(S[2]) - 1.0)[None, None, :]
    u = np.zeros((3,) + S, dtype=np.float32)
    u[0] = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    u[1] = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    G, O = pair(size, q, nu=nu, periodic=True)
    G.init_equilibrium(1.0, u)
    O.init_f(1.0, u.astype(float))
    G.step(100)
    O.run(100)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


def test_step_chunks_equal_single_steps():
    G1, _ = cavity(10, 0)
    G2, _ = cavity(10, 0)
    G1.step(17)
    for _ in range(17):
        G2.step(1)
    s = (slice(None),) + (slice(1, -1),) * 3
    assert np.array_equal(G1.f()[s], G2.f()[s])
    assert np.array_equal(G1.rho()[s[1:]], G2.rho()[s[1:]])


def test_conservation_reduction_matches_oracle():
    G, O = cavity(20, 150, dtype="f64")
    ct = G.cell_types()
    mask = ((ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)).astype(np.uint8)
    a = G.conservation()
    b = O.conservation(mask)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-15), (a, b)


def test_undefined_ghost_link_is_an_error():
    G = asl_amd.Lattice((8, 8, 8), 19, nu=0.1)
    G.init_equilibrium(1.0, (0, 0, 0))
    b = G.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0)  # no ZE
    G.execute_bc(b)
    with pytest.raises(asl_amd.ASLError, match="ASL ERROR"):
        G.step(1)


def test_asl_style_driver_with_deferred_steps():
    """ASL's driver loop through the Python mirror: execute(); BCs; ... read.

    Each execute() is deferred by the library; the host read then sees what
    ASL's per-step rho/u writes leave behind.  Run twice: BCs executed before
    the loop (flow2.cc) and not (raw ghost values in the first step)."""
    import oracle
    for pre in (True, False):
        n, steps, nu = 10, 30, 0.04
        lb = asl_amd.LBGK(asl_amd.Block((n, n, n)), nu, asl_amd.d3q19())
        lb.init()
        asl_amd.LBGKUtilities(lb).initF((0.0, 0.0, 0.0))
        bcs = [asl_amd.generateBCConstantVelocity(lb, (0.05, 0.0, 0.0), ["ZE"]),
               asl_amd.generateBCNoSlip(lb, ["X0", "XE", "Y0", "YE", "Z0"])]
        asl_amd.initAll(bcs)
        if pre:
            asl_amd.executeAll(bcs)
        for _ in range(steps):
            lb.execute()
            asl_amd.executeAll(bcs)
        O = oracle.OracleLattice((n, n, n), 19, nu=nu)
        O.init_f(1.0, (0, 0, 0))
        ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 32, u=(0.05, 0, 0)), O.add_slice_bc(O.NOSLIP, 31)]
        if pre:
            for b in ids:
                O.execute_bc(b)
        O.run(steps, ids)
        s = (slice(1, -1),) * 3
        assert np.array_equal(lb.getRho()[s], O.rho()[s].astype(np.float32)), pre
        assert np.array_equal(lb.getVelocity()[(slice(None),) + s], O.velocity()[(slice(None),) + s].astype(np.float32))


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("compressible", [True, False])
def test_rotating_frame(dtype, compressible):
    """LBGK::setOmega: cavity in a rotating frame, GPU vs oracle bit for bit;
    the term switched on after some plain steps (and off again) as a driver may."""
    G, O = pair((14, 12, 10), 19, nu=0.04, dtype=dtype, compressible=compressible)
    G.init_equilibrium(1.0, (0.01, 0, 0))
    O.init_f(1.0, (0.01, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)),
           add_bc(G, O, _lib.BC_NOSLIP, ALL_BUT_ZE)]
    run(G, O, 20, bcs)
    om = (0.004, -0.002, 0.01)
    G.set_omega(om)
    O.set_omega(om)
    G.step(150)
    O.run(150, [o for _, o in bcs])
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d
    G.set_omega(None)
    O.set_omega(None)
    G.step(30)
    O.run(30, [o for _, o in bcs])
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_periodic_host_views_after_odd_and_even_steps(dtype):
    """f and the FP64 reductions of a periodic lattice after odd steps, when a face
    cell's populations sit in the wrapped neighbour's slots."""
    size = (10, 9, 8)
    rng = np.random.default_rng(3)
    u = (0.03 * rng.standard_normal((3,) + tuple(s + 2 for s in size))).astype(np.float32)
    G, O = pair(size, 19, nu=0.01, dtype=dtype, periodic=True)
    G.init_equilibrium(1.0, u)
    O.init_f(1.0, u.astype(np.float64))
    s = (slice(None),) + (slice(1, -1),) * 3
    for n in (1, 1, 1, 2):
        G.step(n)
        O.run(n)
        assert np.array_equal(G.f()[s], O.f()[s].astype(G.f().dtype))
        mask = np.zeros(G.S, dtype=np.uint8)
        mask[1:-1, 1:-1, 1:-1] = 1
        np.testing.assert_allclose(G.conservation(), O.conservation(mask), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("periodic", [False, True])
def test_write_f_round_trip_and_parity(periodic):
    """aslhip_write_f: the host view reads back bit for bit, and stepping from it
    matches the oracle stepping from the same populations."""
    size = (9, 8, 7)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(9)
    f0 = (1.0 + 0.03 * rng.standard_normal((19,) + S)).astype(np.float32)
    G, O = pair(size, 19, nu=0.03, periodic=periodic)
    G.write_f(f0)
    s = (slice(None),) + (slice(1, -1),) * 3
    assert np.array_equal(G.f(), f0)
    r = np.einsum("i,i...->...", asl_amd.d3q19().w, f0.astype(np.float64))
    assert np.abs(G.rho()[s[1:]] - r[s[1:]]).max() < 1e-6
    O.write_f(f0.astype(np.float64))
    bcs = [] if periodic else [add_bc(G, O, _lib.BC_NOSLIP, 63)]
    run(G, O, 17, bcs)
    assert np.array_equal(G.f()[s], O.f()[s].astype(np.float32))


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_surface_force_and_flux(dtype):
    """ComputeSurfaceForceMap / FluxMap: GPU (post-collision view + resolved
    incoming links) against the oracle's pool state, bit for bit on the interior."""
    size = (20, 14, 12)
    mapv = geometry.normalize(geometry.sphere(size, 3.0, (8.0, 6.5, 5.5)))
    G, O = pair(size, 19, nu=0.04, dtype=dtype)
    G.init_equilibrium(1.0, (0.03, 0, 0))
    O.init_f(1.0, (0.03, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0, 0)),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
           add_bc(G, O, _lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
           add_bc(G, O, _lib.BC_NOSLIP_MAP, mapv=mapv)]
    run(G, O, 31, bcs)
    inner = (slice(1, -1),) * 3
    fg, fo = G.surface_force(mapv), O.surface_force(mapv)
    qg, qo = G.surface_flux(mapv), O.surface_flux(mapv)
    assert np.abs(fo[(slice(None),) + inner]).max() > 1e-4
    assert np.array_equal(fg[(slice(None),) + inner], fo[(slice(None),) + inner].astype(fg.dtype))
    assert np.array_equal(qg[inner], qo[inner].astype(qg.dtype))


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("compressible", [True, False])
def test_turbulence_periodic(q, dtype, compressible):
    """LBGKTurbulence on a periodic lattice started far from equilibrium (cells
    in all three regimes of the limiter), GPU vs oracle bit for bit."""
    size = (12, 10, 9)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(q)
    f0 = 1.0 + rng.uniform(0.01, 1.6, S) * rng.standard_normal((q,) + S)
    G, O = pair(size, q, nu=0.2, dtype=dtype, periodic=True, compressible=compressible)
    G.set_turbulence(True)
    O.set_turbulence(True)
    G.write_f(f0)
    O.write_f(G.f().astype(np.float64))
    s = (slice(None),) + (slice(1, -1),) * 3
    for n in (1, 24):
        G.step(n)
        O.run(n)
        fo = O.f()[s].astype(G.f().dtype)
        assert np.isfinite(fo).all()
        assert np.array_equal(G.f()[s], fo)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_turbulence_sphere_channel_d3q15(dtype):
    """The locomotive set-up in miniature: LBGKTurbulence, D3Q15, velocity map
    tunnel + no-slip body + pressure-velocity inlet/outlet, bit for bit."""
    size = (24, 14, 12)
    body = geometry.normalize(geometry.sphere(size, 3.0, (9.0, 6.5, 5.5)))
    G, O = pair(size, 15, nu=0.001, dtype=dtype)
    G.set_turbulence(True)
    O.set_turbulence(True)
    G.init_equilibrium(1.0, (0.1, 0, 0))
    O.init_f(1.0, (0.1, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_NOSLIP_MAP, mapv=body),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE_VELOCITY, _lib.X0 | _lib.XE, p=1.0, u=(0.1, 0, 0)),
           add_bc(G, O, _lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE)]
    run(G, O, 60, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d
    G.set_turbulence(False)
    O.set_turbulence(False)
    G.step(10)
    O.run(10, [o for _, o in bcs])
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


def test_turbulence_and_closure_exclude_each_other():
    G = asl_amd.Lattice((8, 8, 8), 19, periodic=True)
    G.set_turbulence(True)
    with pytest.raises(asl_amd.ASLError):
        G.set_closure(np.zeros(_lib.CLOSURE_NPARAM, dtype=np.float32))


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_abd_checkpoint_restart(tmp_path, dtype):
    """Populations saved as an ABD file after some steps and loaded into a fresh
    lattice with the same BCs, executed after the load as a driver does after
    initF: the restarted run equals the uninterrupted one bit for bit."""
    size = (20, 14, 12)
    mapv = geometry.normalize(geometry.sphere(size, 3.0, (8.0, 6.5, 5.5)))

    def make(restore=None):
        L = asl_amd.Lattice(size, 19, dtype=dtype, nu=0.04)
        if restore:
            L.load_abd(restore)
        else:
            L.init_equilibrium(1.0, (0.03, 0, 0))
        bcs = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0, 0)),
               L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
               L.add_slice_bc(_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
               L.add_map_bc(_lib.BC_NOSLIP_MAP, mapv)]
        for b in bcs:
            L.execute_bc(b)
        return L

    A = make()
    A.step(37)
    A.save_abd(tmp_path / "f.abd")
    A.step(20)
    B = make(tmp_path / "f.abd")
    B.step(20)
    s = (slice(None),) + (slice(1, -1),) * 3
    ct = A.cell_types()
    comp = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    assert np.array_equal(A.f()[s][:, comp[s[1:]]], B.f()[s][:, comp[s[1:]]])
    assert np.array_equal(A.rho()[comp], B.rho()[comp])
