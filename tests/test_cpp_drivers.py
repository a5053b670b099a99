"""The C++ host side: ASL-style drivers over the asl:: compatibility layer.

GPU: the C++ parity programs (tests/cpp/*.cc, built by build() into
asl_amd/bin) run the lid-driven cavity and the sphere channel through
asl::LBGK / asl::generateBC* exactly as ASL's examples/flow drivers do, and
compare with the CPU oracle; they exit with EXIT_SUCCESS like ASL's tests.

CPU: ASL's own flow drivers compile and link unchanged against the
compatibility headers (skipped where the ASL source tree is not mounted).
"""
import os
import subprocess

import numpy as np

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "asl_amd", "bin")
REF_FLOW = "/root/reference/examples/flow"


def _run(name, *args, timeout=600):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "%s not built (run __graft_entry__.build())" % exe
    p = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(p.stdout, p.stderr)
    return p


@pytest.mark.gpu
def test_cpp_cavity_64_1000():
    p = _run("cavity_parity", 64, 1000)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout
    assert "device velocity view equals host view: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_sphere_channel():
    p = _run("sphere_parity", 64, 32, 300)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout and "cell types match masks: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_sphere_channel_full_size():
    # the benchmark lattice (256x128x128, sphere r = 16 at (64,64,64)), 20 steps
    p = _run("sphere_parity", 256, 128, 20, timeout=900)
    assert p.returncode == 0
    assert "bit-identical: yes" in p.stdout and "cell types match masks: yes" in p.stdout


@pytest.mark.gpu
def test_cpp_stl_body_turbulence(tmp_path):
    # locomotive-style: STL body via readSurface(file, Block), LBGKTurbulence D3Q15,
    # tunnel velocity map, visualisation BCs, surface force, .vti output
    from test_gpu_surface_map import icosphere
    from test_oracle_surface_map import write_stl_binary
    path = tmp_path / "body.stl"
    write_stl_binary(path, icosphere(0.8, (2.6, 1.5, 3.0)))  # file (x, y, z) = ASL (z, y, x)
    p = _run("stl_body_parity", path, 200, tmp_path / "out")
    assert p.returncode == 0
    assert "map bit-identical: yes" in p.stdout and ", bit-identical: yes" in p.stdout
    assert (tmp_path / "out_0.vti").stat().st_size > 10000


@pytest.mark.gpu
def test_cpp_vtp_body_equals_stl_body(tmp_path):
    # readSurface(file.vtp, Block) of the same triangles (appended, base64,
    # zlib: what VTK's writer produces by default) gives the STL run's outputs
    from test_gpu_surface_map import icosphere
    from test_oracle_surface_map import write_stl_binary
    from vtp import write_vtp
    tri = np.asarray(icosphere(0.8, (2.6, 1.5, 3.0)), dtype=np.float32)
    write_stl_binary(tmp_path / "body.stl", tri)
    write_vtp(tmp_path / "body.vtp", [{"points": tri.reshape(-1, 3),
                                       "polys": [[3 * k, 3 * k + 1, 3 * k + 2] for k in range(len(tri))]}],
              form="appended", appended_encoding="base64", compress=True, block=32768)
    outs = []
    for ext in ("stl", "vtp"):
        p = _run("stl_body_parity", tmp_path / ("body." + ext), 60, tmp_path / ext)
        assert p.returncode == 0
        assert "map bit-identical: yes" in p.stdout and ", bit-identical: yes" in p.stdout
        outs.append((tmp_path / (ext + "_0.vti")).read_bytes())
    assert outs[0] == outs[1]


@pytest.mark.skipif(not os.path.isdir(REF_FLOW), reason="ASL source tree not mounted")
@pytest.mark.parametrize("driver", ["flow", "flow2", "flow3", "flowRotatingCylinders", "locomotive",
                                    "locomotive_laminar", "locomotive_stability", "bus_wind", "compressor"])
def test_asl_flow_drivers_compile_unchanged(driver, tmp_path):
    cmd = ["g++", "-std=c++17", "-O0", "-Wno-ignored-qualifiers", "-I", os.path.join(ROOT, "asl_amd", "compat", "include"),
           "-I", os.path.join(ROOT, "include"), os.path.join(REF_FLOW, driver + ".cc"), "-o", str(tmp_path / driver),
           "-L", os.path.join(ROOT, "asl_amd"), "-laslcompat", "-laslhip"]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]


def _compare_driver_output(vti_path, size, q, nu, bcs, prog, steps):
    """The driver's output against the oracle run of the same set-up.  The map
    comes from the driver's own geometry rebuilt as a distance-function program
    (the written map covers the internal cells only; the ghost layer matters
    for the BCs) and must equal the written map bit for bit first."""
    import numpy as np

    import oracle
    from vti import read_vti
    out = read_vti(vti_path)
    S = tuple(s + 2 for s in size)
    full = oracle.df_map(prog, S, (-1.0, -1.0, -1.0), 1.0)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(full[inner], out["map"])
    O = oracle.OracleLattice(size, q, fp64=False, nu=nu)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = []
    for kind, faces, p, u in bcs:
        ids.append(O.add_map_bc(O.NOSLIP_MAP, full) if kind == "map" else O.add_slice_bc(kind, faces, p=p, u=u))
    for b in ids:
        O.execute_bc(b)
    O.run(steps, ids)
    fluid = out["map"] > 0
    rho_o = O.rho()[inner].astype(np.float32)
    v_o = O.velocity()[(slice(None),) + inner].astype(np.float32)
    assert fluid.sum() > 1000
    assert np.array_equal(out["rho"][fluid], rho_o[fluid])
    assert np.array_equal(out["v"][:, fluid], v_o[:, fluid])
    return np.abs(v_o[:, fluid]).max()


@pytest.mark.gpu
def test_asl_flow2_driver_unchanged(tmp_path):
    """ASL's examples/flow/flow2.cc, built unchanged against the compat layer,
    run on the GPU: its 11th output (901 steps) equals the oracle run of the
    same set-up bit for bit on the fluid cells."""
    from asl_driver_geometry import flow2_mirrors
    from oracle import OracleLattice as OL
    exe = os.path.join(BIN, "asl_flow2")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    umax = _compare_driver_output(tmp_path / "flow2_10.vti", (200, 200, 25), 15, float(np.float32(0.0125)),
                                  [(OL.CONSTANT_PRESSURE, 1 | 2, 1.0, None),
                                   (OL.CONSTANT_VELOCITY, 32, 0.0, (0.1, 0.0, 0.0)),
                                   (OL.NOSLIP, 4 | 8 | 16, 0.0, None), ("map", None, 0.0, None)],
                                  flow2_mirrors(), 901)
    assert umax > 0.01


@pytest.mark.gpu
def test_asl_flow3_driver_unchanged(tmp_path):
    """ASL's examples/flow/flow3.cc (pressure-driven flow through a cylinder
    array, BCConstantPressure objects with addSliceX0/XE) likewise."""
    from asl_driver_geometry import flow3_cylinders
    from oracle import OracleLattice as OL
    exe = os.path.join(BIN, "asl_flow3")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    umax = _compare_driver_output(tmp_path / "flow3Res_10.vti", (50, 50, 50), 15, float(np.float32(0.00625)),
                                  [(OL.CONSTANT_PRESSURE, 1, 1.2, None), (OL.CONSTANT_PRESSURE, 2, 0.8, None),
                                   (OL.NOSLIP, 4 | 8 | 16 | 32, 0.0, None), ("map", None, 0.0, None)],
                                  flow3_cylinders(), 901)
    assert umax > 0.01


@pytest.mark.gpu
def test_asl_flow_rotating_cylinders_driver_unchanged(tmp_path):
    """ASL's examples/flow/flowRotatingCylinders.cc (Couette flow between
    counter-rotating cylinders: two velocity-map BCs with rotation fields and
    a computational-domain map), built unchanged: its second output (1001
    steps) equals the oracle bit for bit on the computational domain."""
    from asl_driver_geometry import rotating_cylinders
    from vti import read_vti

    import oracle
    exe = os.path.join(BIN, "asl_flowRotatingCylinders")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = read_vti(tmp_path / "flowRotCylRes_2.vti")
    size, w = (100, 100, 150), 2e-3
    ex, inner_cyl, dom, c = rotating_cylinders(size)
    S = tuple(s + 2 for s in size)
    maps = [oracle.df_map(prg, S, (-1.0, -1.0, -1.0), 1.0) for prg in (ex, inner_cyl, dom)]
    inner = (slice(1, -1),) * 3
    assert np.array_equal(maps[2][inner], out["map"])
    O = oracle.OracleLattice(size, 15, fp64=False, nu=float(np.float32(0.01)))
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_map_bc(O.VELOCITY_MAP, maps[0], cdomain=maps[2], u=(0, 0, 0), rot=(0, 0, w) + tuple(c)),
           O.add_map_bc(O.VELOCITY_MAP, maps[1], cdomain=maps[2], u=(0, 0, 0), rot=(0, 0, -2 * w) + tuple(c)),
           O.add_slice_bc(O.NOSLIP, 16 | 32)]
    for b in ids:
        O.execute_bc(b)
    O.run(1001, ids)
    fluid = (maps[0][inner] > 0) & (maps[1][inner] > 0) & (out["map"] > 0)
    v_o = O.velocity()[(slice(None),) + inner].astype(np.float32)
    assert fluid.sum() > 100000 and np.abs(v_o[:, fluid]).max() > 0.01
    assert np.array_equal(out["rho"][fluid], O.rho()[inner].astype(np.float32)[fluid])
    assert np.array_equal(out["v"][:, fluid], v_o[:, fluid])


def _locomotive_case(tmp_path, name, args, box, extent, gpos, dx, rdiv, nu, turbulence, steps, out_index, mapname):
    """Runs one of ASL's locomotive drivers, built unchanged, on a box-shaped STL
    body (box: lo, hi in file coordinates (x, y, z) = ASL (z, y, x)) and checks
    its output number out_index against the oracle run of the same set-up: the
    body map (readSurface(file, Block) on the extent / dx grid at gpos), the
    tunnel level set, ρ and u on the fluid nodes and the surface force on the
    interior, all bit for bit."""
    from asl_driver_geometry import locomotive_tunnel
    from test_oracle_surface_map import cube_triangles, write_stl_binary
    from vti import read_vti

    import oracle
    exe = os.path.join(BIN, "asl_" + name)
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    tri = cube_triangles(np.array(box[0]), np.array(box[1]))
    write_stl_binary(tmp_path / "body.stl", tri)
    p = subprocess.run([exe, "--input", str(tmp_path / "body.stl")] + args,
                       capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = read_vti(tmp_path / ("%s_%d.vti" % (name, out_index)))
    grid = [int(v * (1.0 / dx)) for v in extent]                    # the Block given to readSurface
    size = tuple(s - 2 for s in grid)                               # the lattice's internal block
    ipos = [v + dx for v in gpos]                                   # offset(Block, -1)
    tpos = [v - dx for v in ipos]                                   # the tunnel container: offset(block, 1)
    t_asl = tri.astype(np.float32).astype(np.float64)[..., ::-1]
    h = [((gpos[d] + (grid[d] - 1) * dx) - gpos[d]) / (grid[d] - 1) for d in range(3)]
    body = oracle.surface_distance_map(t_asl, grid, gpos, h, 3.6 * dx, 1.2 * dx, dx, 1.8)
    tunnel = oracle.df_map(locomotive_tunnel(ipos, size, dx, rdiv), grid, tpos, dx)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(body[inner], out[mapname]) and np.array_equal(tunnel[inner], out["tunnel"])
    O = oracle.OracleLattice(size, 15, fp64=False, nu=nu)
    O.set_turbulence(turbulence)
    O.init_f(1.0, (0.1, 0.0, 0.0))
    ids = [O.add_map_bc(O.VELOCITY_MAP, tunnel, u=(0.1, 0.0, 0.0)), O.add_map_bc(O.NOSLIP_MAP, body),
           O.add_slice_bc(O.CONSTANT_PRESSURE_VELOCITY, 1 | 2, p=1.0, u=(0.1, 0.0, 0.0))]
    for b in ids:
        O.execute_bc(b)
    O.run(steps, ids)
    fluid = (body[inner] > 0) & (tunnel[inner] > 0)
    assert fluid.sum() > 100000
    assert np.array_equal(out["rho"][fluid], O.rho()[inner].astype(np.float32)[fluid])
    assert np.array_equal(out["v"][:, fluid], O.velocity()[(slice(None),) + inner].astype(np.float32)[:, fluid])
    fo = O.surface_force(body)[(slice(None),) + inner].astype(np.float32)
    assert np.abs(fo).max() > 1e-3
    assert np.array_equal(out["force"], fo)


@pytest.mark.gpu
def test_asl_locomotive_driver_unchanged(tmp_path):
    """ASL's examples/flow/locomotive.cc built unchanged: an STL body read with
    readSurface(file, Block), the tunnel level set, LBGKTurbulence on D3Q15,
    the velocity-map tunnel, visualisation BCs and the surface force; run on
    the GPU with a 10 m box body and --dx 0.25 --nu 0.01 --iterations 1001
    (the driver hands ν to the method in lattice units: with its default
    0.001 a 0.16 m lattice diverges within 1000 steps in the oracle), its
    output after 1000 steps equals the oracle bit for bit."""
    _locomotive_case(tmp_path, "locomotive", ["--dx", "0.25", "--nu", "0.01", "--iterations", "1001"],
                     ([7.5, 9.8, -15.0], [10.5, 11.8, -5.0]), (40.0, 10.0, 15.0), [-30.0, 8.58, 1.53],
                     float(np.float32(0.25)),  # asl::Parameter<float> dx, widened to the Block's double
                     2.1, float(np.float32(0.01)), True, 1000, 1, "map")


@pytest.mark.gpu
def test_asl_locomotive_laminar_driver_unchanged(tmp_path):
    """ASL's examples/flow/locomotive_laminar.cc built unchanged (plain LBGK
    D3Q15, dx 0.5 m, ν = 0.01 m²/s → 0.04 in lattice units, the tunnel radius
    half the block height): with a 30 m box body its first output after the
    start (200 steps) equals the oracle bit for bit."""
    _locomotive_case(tmp_path, "locomotive_laminar", [], ([-7.0, 0.5, 10.0], [-3.0, 4.5, 40.0]),
                     (150.0, 37.0, 53.2), [-10.0, -1.63, -31.6], 0.5, 2.0, float(np.float32(0.01 / 0.5 / 0.5)),
                     False, 200, 1, "locomotive")



def _surface_box_variant(tri, dx, offs):
    """readSurface(file, dx, offset_X0, offset_XE, offset_Y0, offset_YE,
    offset_Z0, offset_ZE) of the compat layer on a triangle set given in file
    coordinates: the adjusted model bounds (ASL axis order) grown by the
    offsets, sample counts and spacings, and the lattice's internal block
    (offset(grid, -1), dx the spacing along ASL's last axis)."""
    t = tri.astype(np.float32).astype(np.float64)[..., ::-1]
    pts = t.reshape(-1, 3)
    lo, hi = pts.min(0), pts.max(0)
    m = (hi - lo).max()
    lo, hi = lo - m * 0.0125, hi + m * 0.0125
    o0, o1 = np.array(offs[0::2]), np.array(offs[1::2])
    l, u = lo - (hi - lo) * o0, hi + (hi - lo) * o1
    grid = [int((u[d] - l[d]) / dx) for d in range(3)]
    h = [(u[d] - l[d]) / (grid[d] - 1) for d in range(3)]
    return t, grid, list(l), h


def _force_nodes(body):
    """Interior nodes whose surface force ASL defines.  readSurface's dx
    variants put a body with offset 0 on the block's face (bus_wind and
    locomotive_stability: Y0), so its shell meets the outer ghost layer: a
    link from a boundary node (map <= 0) into an outer-layer node with map > 0
    carries what ASL leaves in memory that no step or BC writes (NaN in the
    oracle), and those nodes are left out."""
    import oracle
    outer = np.ones(body.shape, dtype=bool)
    outer[(slice(1, -1),) * 3] = False
    risky = np.zeros(body.shape, dtype=bool)
    c, _, _ = oracle.template(15)
    for ci in c[1:]:
        risky |= np.roll(outer & (body > 0), tuple(-ci), axis=(0, 1, 2))
    risky &= body <= 0
    return ~risky[(slice(1, -1),) * 3]


def _body_and_block(tri, dx, offs):
    import oracle
    t, grid, gpos, h = _surface_box_variant(tri, dx, offs)
    body = oracle.surface_distance_map(t, grid, gpos, h, 3.6 * dx, 1.8 * dx, dx, 1.8)
    bdx = h[2]
    size = tuple(s - 2 for s in grid)
    ipos = [v + bdx for v in gpos]  # offset(grid, -1)
    return body, grid, size, ipos, bdx


@pytest.mark.gpu
def test_asl_flow_driver_unchanged(tmp_path):
    """ASL's examples/flow/flow.cc built unchanged (D3Q15, 300x50x50, pressure
    drop along x past an infinite cylinder, no-slip walls, the visualisation
    fill of u): it runs its 1000 steps on the GPU; its one output (written
    before the loop, after the BCs and the fill) equals the oracle's."""
    from asl_driver_geometry import cylinder_inf
    from vti import read_vti

    import oracle
    exe = os.path.join(BIN, "asl_flow")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = read_vti(tmp_path / "flow_0.vti")
    size = (300, 50, 50)
    S = tuple(s + 2 for s in size)
    ball = oracle.df_map(cylinder_inf(.125 * 50, [0., 1., 0.], [.45 * 50, .45 * 50, .45 * 50]), S,
                         (-1.0, -1.0, -1.0), 1.0)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(ball[inner], out["map"])
    O = oracle.OracleLattice(size, 15, fp64=False, nu=float(np.float32(0.00625)))
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_PRESSURE, 1, p=1.2), O.add_slice_bc(O.CONSTANT_PRESSURE, 2, p=0.8),
           O.add_slice_bc(O.NOSLIP, 4 | 8 | 16 | 32), O.add_map_bc(O.NOSLIP_MAP, ball)]
    for b in ids:
        O.execute_bc(b)
    O.vis_fill(1, ball)
    guard = (ball[inner] <= 0) & (ball[inner] > -0.9999)
    keep = (ball[inner] > 0) | guard
    assert guard.sum() > 1000
    assert np.array_equal(out["rho"][ball[inner] > 0], O.rho()[inner].astype(np.float32)[ball[inner] > 0])
    assert np.array_equal(out["v"][:, keep], O.velocity()[(slice(None),) + inner].astype(np.float32)[:, keep])


@pytest.mark.gpu
def test_asl_bus_wind_driver_unchanged(tmp_path):
    """ASL's examples/flow/bus_wind.cc built unchanged on a box-shaped STL
    body (readSurface(file, dx 8, offsets), D3Q15, pressure-velocity on all six
    faces, the visualisation fill and the surface force): it runs its 100
    steps on the GPU; its one output (written before the loop) equals the
    oracle's, body map, rho, u (fill included) and force bit for bit."""
    from test_oracle_surface_map import cube_triangles, write_stl_binary
    from vti import read_vti

    import oracle
    exe = os.path.join(BIN, "asl_bus_wind")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    tri = cube_triangles(np.array([0.0, 0.0, 0.0]), np.array([400.0, 100.0, 120.0]))
    write_stl_binary(tmp_path / "bus.stl", tri)
    p = subprocess.run([exe, "--input", str(tmp_path / "bus.stl")], capture_output=True, text=True, timeout=600,
                       cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = read_vti(tmp_path / "bus_wind_0.vti")
    body, grid, size, _, _ = _body_and_block(tri, 8.0, (1.5, .25, 0., 1., 3., 1.))
    inner = (slice(1, -1),) * 3
    assert np.array_equal(body[inner], out["bus"])
    U = (-0.1, 0.0, -0.05)
    O = oracle.OracleLattice(size, 15, fp64=False, nu=float(np.float32(0.01)))
    O.init_f(1.0, U)
    ids = [O.add_map_bc(O.NOSLIP_MAP, body), O.add_slice_bc(O.CONSTANT_PRESSURE_VELOCITY, 63, p=1.0, u=U)]
    for b in ids:
        O.execute_bc(b)
    O.vis_fill(1, body)
    guard = (body[inner] <= 0) & (body[inner] > -0.9999)
    keep = (body[inner] > 0) | guard
    assert guard.sum() > 1000
    assert np.array_equal(out["rho"][body[inner] > 0], O.rho()[inner].astype(np.float32)[body[inner] > 0])
    assert np.array_equal(out["v"][:, keep], O.velocity()[(slice(None),) + inner].astype(np.float32)[:, keep])
    # the mirrored stream (0 where a boundary node has no fluid neighbour)
    assert (out["v"][0][guard] >= 0).all() and (out["v"][0][guard] > 0).mean() > 0.5
    ok = _force_nodes(body)
    fo = O.surface_force(body)[(slice(None),) + inner].astype(np.float32)
    assert ok.sum() > 0.99 * ok.size and np.abs(fo[:, ok]).max() > 1e-4
    assert np.array_equal(out["force"][:, ok], fo[:, ok])


@pytest.mark.gpu
def test_asl_locomotive_stability_driver_unchanged(tmp_path):
    """ASL's examples/flow/locomotive_stability.cc built unchanged on a
    box-shaped STL body placed under the driver's four nozzles
    (readSurface(file, dx 0.25, offsets), D3Q15, pressure-velocity on all six
    faces and on the nozzle map, the visualisation fill, the surface force):
    it runs its 40,000 steps on the GPU; its output after 1000 steps equals
    the oracle bit for bit (maps, rho and u on the fluid nodes, force)."""
    from asl_driver_geometry import cylinder
    from test_oracle_surface_map import cube_triangles, write_stl_binary
    from vti import read_vti

    from asl_amd import geometry as g
    import oracle
    exe = os.path.join(BIN, "asl_locomotive_stability")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    # file (x, y, z) = ASL (z, y, x): the body spans ASL x in [-16, -6], y in [7, 9], z in [3, 4]
    tri = cube_triangles(np.array([3.0, 7.0, -16.0]), np.array([4.0, 9.0, -6.0]))
    write_stl_binary(tmp_path / "body.stl", tri)
    p = subprocess.run([exe, "--input", str(tmp_path / "body.stl")], capture_output=True, text=True,
                       timeout=900, cwd=tmp_path)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    out = read_vti(tmp_path / "locomotive_stability_1.vti")
    dx = float(np.float32(0.25))  # asl::Parameter<float> dx, widened to double
    body, grid, size, ipos, bdx = _body_and_block(tri, dx, (.5, 1., 0., 1., 2., 4.))
    inner = (slice(1, -1),) * 3
    assert np.array_equal(body[inner], out["locomotive"])
    # generateNozzels(block): four cylinders r 0.3, height 1 along z, normalised by the block's dx
    noz = None
    for i in range(4):
        c = cylinder(.3, [0., 0., 1.], [-20. + i * (-6. + 20.) / 4., 10.33, 3.08 + .5])
        noz = c if noz is None else g.df_or(noz, c)
    tpos = [v - bdx for v in ipos]
    nozzles = oracle.df_map(g.df_normalize(noz, bdx), grid, tpos, bdx)
    assert np.array_equal(nozzles[inner], out["nozzels"])
    assert (nozzles[inner] <= 0).sum() > 20
    O = oracle.OracleLattice(size, 15, fp64=False, nu=float(np.float32(0.01)))
    O.init_f(1.0, (0.1, 0.0, 0.05))
    ids = [O.add_map_bc(O.NOSLIP_MAP, body),
           O.add_slice_bc(O.CONSTANT_PRESSURE_VELOCITY, 63, p=1.0, u=(0.1, 0.0, 0.05)),
           O.add_map_bc(O.PRESSURE_VELOCITY_MAP, nozzles, p=1.0, u=(0.0, 0.0, -0.1))]
    for b in ids:
        O.execute_bc(b)
    O.run(1000, ids)
    fluid = (body[inner] > 0) & (nozzles[inner] > 0)
    assert fluid.sum() > 20000
    v_o = O.velocity()[(slice(None),) + inner].astype(np.float32)
    assert np.abs(v_o[:, fluid]).max() > 0.05
    assert np.array_equal(out["rho"][fluid], O.rho()[inner].astype(np.float32)[fluid])
    assert np.array_equal(out["v"][:, fluid], v_o[:, fluid])
    ok = _force_nodes(body)
    fo = O.surface_force(body)[(slice(None),) + inner].astype(np.float32)
    assert ok.sum() > 0.99 * ok.size and np.abs(fo[:, ok]).max() > 1e-4
    assert np.array_equal(out["force"][:, ok], fo[:, ok])


@pytest.mark.gpu
def test_asl_compressor_driver_with_outer_average_opt_in(tmp_path):
    """ASL's examples/flow/compressor.cc built unchanged (D3Q15, rotating
    frame, rotating velocity-map casing whose boundary nodes in the first and
    last z layer average the density over outer-layer cells, pressure on Z0 /
    ZE, no-slip rotor from readSurface(file, Block), visualisation fills) on a
    box-shaped rotor.  Without the opt-in it stops with ASLHIP_EUNDEF; with
    ASLHIP_VMAP_SKIP_OUTER=1 (those neighbours left out of the average, the
    rule tests/test_gpu_parity.py checks bit for bit against the oracle) it
    runs its 10,000 steps and writes finite fields."""
    from test_oracle_surface_map import cube_triangles, write_stl_binary
    from vti import read_vti
    exe = os.path.join(BIN, "asl_compressor")
    if not os.path.exists(exe):
        pytest.skip("ASL's flow drivers were not built (ASL source tree not mounted at build time)")
    # file (x, y, z) = ASL axes (2, 1, 0): the rotor inside the casing
    write_stl_binary(tmp_path / "rotor.stl", cube_triangles(np.array([-55.0, -14.0, -5.0]),
                                                            np.array([-35.0, -5.0, 5.0])))
    env = {k: v for k, v in os.environ.items() if k != "ASLHIP_VMAP_SKIP_OUTER"}
    p = subprocess.run([exe, "--input", "rotor.stl"], capture_output=True, text=True, timeout=600, cwd=tmp_path,
                       env=env)
    assert p.returncode != 0 and "averages a neighbour that is not a computed cell" in (p.stdout + p.stderr)
    env["ASLHIP_VMAP_SKIP_OUTER"] = "1"
    p = subprocess.run([exe, "--input", "rotor.stl"], capture_output=True, text=True, timeout=600, cwd=tmp_path,
                       env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "Finished" in p.stdout
    last = read_vti(tmp_path / "compressor_5.vti")  # after 10,000 steps
    fluid = (last["case"] > 0) & (last["compressor"] > 0)
    assert fluid.sum() > 100000 and (last["compressor"] <= 0).sum() > 1000
    assert np.isfinite(last["rho"][fluid]).all() and np.isfinite(last["v"][:, fluid]).all()
    # the output after 2000 steps against the oracle run of the same set-up
    # (with the same opt-in): maps, rho on the fluid nodes, u on the fluid
    # and the visualisation-filled boundary nodes, bit for bit
    import oracle
    from asl_amd import geometry
    from asl_driver_geometry import block_bpos, cylinder_inf
    out = read_vti(tmp_path / "compressor_1.vti")
    dx = 0.5
    grid = [int(v * (1.0 / dx)) for v in (33.0, 33.0, 90.0)]       # the Block given to readSurface
    gpos = [-16.65, -26.15, -67.7]
    size = tuple(n - 2 for n in grid)                               # the lattice's internal block
    ipos = [v + dx for v in gpos]
    tri = cube_triangles(np.array([-55.0, -14.0, -5.0]), np.array([-35.0, -5.0, 5.0]))
    t_asl = tri.astype(np.float32).astype(np.float64)[..., ::-1]
    h = [((gpos[d] + (grid[d] - 1) * dx) - gpos[d]) / (grid[d] - 1) for d in range(3)]
    body = oracle.surface_distance_map(t_asl, grid, gpos, h, 3.6 * dx, 1.2 * dx, dx, 1.8)
    bp = block_bpos(ipos, size, dx)
    r_case = (bp[1] - ipos[1]) / 2.
    centre = [.5 * (bp[d] + ipos[d]) for d in range(3)]
    centre[1] = ipos[1] + r_case
    case = oracle.df_map(geometry.df_normalize(geometry.df_neg(cylinder_inf(r_case, [0., 0., 1.], centre)), dx),
                         grid, gpos, dx)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(body[inner], out["compressor"]) and np.array_equal(case[inner], out["case"])
    w_num = 6. * 3.14 * 2. / 60. * .001
    rot = (0., 0., w_num / dx) + tuple(.5 * (bp[d] + ipos[d]) for d in range(3))
    O = oracle.OracleLattice(size, 15, fp64=False, nu=float(np.float32(0.2)))
    O.set_geometry(dx, ipos)
    O.set_omega((0., 0., w_num))
    O.set_vmap_skip_outer(True)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_map_bc(O.VELOCITY_MAP, case, u=(0., 0., 0.), rot=rot),
           O.add_slice_bc(O.CONSTANT_PRESSURE, 16 | 32, p=1.0), O.add_map_bc(O.NOSLIP_MAP, body)]

    def vis():
        O.vis_fill(3, case, u=(0., 0., 0.), rot=rot)
        O.vis_fill(1, body)
    for b in ids:
        O.execute_bc(b)
    vis()
    O.run(2000, ids)
    vis()
    guard = lambda m: (m[inner] <= 0) & (m[inner] > -0.9999)  # noqa: E731
    fluid = (case[inner] > 0) & (body[inner] > 0)
    keep = fluid | (guard(case) & (body[inner] > 0)) | (guard(body) & (case[inner] > 0))
    assert np.array_equal(out["rho"][fluid], O.rho()[inner].astype(np.float32)[fluid])
    v_o = O.velocity()[(slice(None),) + inner].astype(np.float32)
    assert np.array_equal(out["v"][:, keep], v_o[:, keep])
    assert np.abs(v_o[:, fluid]).max() > 1e-4

