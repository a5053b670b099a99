"""readSurface's distance map on the GPU (aslhip_surface_distance_map) against
the oracle, bit for bit: the same double-precision point-triangle distance,
rounded to float, minimum over triangles, capped, rescaled in float.

Meshes: a subdivided icosahedron written as a binary STL and read back
through the C ABI (the locomotive drivers' path, readSurface(file, Block));
a random triangle soup with degenerate and far-away triangles; one triangle
larger than the grid (many workgroup pieces).  The map then drives a
no-slip map BC on a lattice, GPU vs oracle.
"""
import numpy as np
import pytest

import oracle
from asl_amd import _lib, geometry
from gpu_common import add_bc, compare, pair, run
from test_oracle_surface_map import write_stl_binary

pytestmark = pytest.mark.gpu


def icosphere(radius, centre, levels=2):
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    tri = [np.array([v[a], v[b], v[c]]) for a, b, c in f]
    for _ in range(levels):
        out = []
        for a, b, c in tri:
            ab, bc, ca = [(x + y) / np.linalg.norm(x + y) for x, y in ((a, b), (b, c), (c, a))]
            out += [np.array(q) for q in ((a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca))]
        tri = out
    return np.array(tri) * radius + np.asarray(centre, dtype=np.float64)


def both(tri, size, pos, h, cap, shift, div1, div2):
    g = geometry.surface_map(tri, size, pos, h, cap, shift, div1, div2)
    o = oracle.surface_distance_map(tri, size, pos, h, cap, shift, div1, div2)
    return g, o


def test_stl_sphere_block_variant(tmp_path):
    size, dx, position = (30, 26, 22), 0.5, (-1.0, 2.0, 0.5)
    tri = icosphere(4.0, (6.0, 7.5, 6.0))  # file coordinates (x, y, z)
    path = tmp_path / "sphere.stl"
    write_stl_binary(path, tri)
    got = geometry.read_surface(path, size, dx, position)
    tri_asl = geometry.stl_to_asl_axes(tri.astype(np.float32).astype(np.float64))
    pos, h = geometry.block_nodes(size, dx, position)
    ref = oracle.surface_distance_map(tri_asl, size, pos, h, 3.6 * dx, 1.2 * dx, dx, 1.8)
    assert np.array_equal(got, ref)
    shell = got <= 0
    assert 0.02 * got.size < shell.sum() < 0.5 * got.size
    assert got.max() == np.float32(((np.float32(np.sqrt(np.float32((3.6 * dx) ** 2))) - np.float32(1.2 * dx))
                                    / np.float32(dx)) / np.float32(1.8))


def test_triangle_soup_with_degenerate_and_distant_triangles():
    rng = np.random.default_rng(12)
    size = (24, 20, 28)
    tri = rng.uniform(-4, 30, size=(300, 3, 3))
    tri[::17, 2] = tri[::17, 0]                                   # repeated vertex
    tri[5::17, 2] = tri[5::17, 0] + 0.5 * (tri[5::17, 1] - tri[5::17, 0])  # collinear
    tri[9::23] += 100.0                                           # outside the grid
    g, o = both(tri, size, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2.5, 0.9, 1.0, 1.8)
    assert np.array_equal(g, o)


def test_large_triangle_spans_many_pieces():
    size = (64, 60, 70)
    tri = np.array([[[-10.0, -5.0, 30.3], [120.0, 2.0, 35.1], [5.0, 130.0, 33.7]],
                    [[31.2, -3.0, -4.0], [33.9, 90.0, 8.0], [32.4, 7.0, 95.0]]])
    g, o = both(tri, size, (0.25, -0.5, 0.0), (1.0, 1.0, 1.0), 3.0, 1.2, 1.0, 1.8)
    assert np.array_equal(g, o)
    assert (g <= 0).sum() > 1000


def test_empty_mesh_is_the_cap():
    g = geometry.surface_map(np.zeros((0, 3, 3)), (5, 6, 7), (0, 0, 0), (1, 1, 1), 2.0, 0.0, 1.0, 1.0)
    assert (g == np.float32(2.0)).all()


def test_bad_arguments_raise():
    import asl_amd
    with pytest.raises(asl_amd.ASLError):
        geometry.surface_map(np.zeros((1, 3, 3)), (5, 6, 7), (0, 0, 0), (1, 0, 1), 2.0, 0.0, 1.0, 1.0)
    with pytest.raises(asl_amd.ASLError):
        geometry.surface_map(np.full((1, 3, 3), np.nan), (5, 6, 7), (0, 0, 0), (1, 1, 1), 2.0, 0.0, 1.0, 1.0)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_stl_body_in_a_channel(tmp_path, dtype):
    """The STL map as the obstacle of a channel (no-slip map BC), GPU vs oracle."""
    size = (26, 14, 14)
    S = tuple(s + 2 for s in size)
    tri = icosphere(3.2, (7.0, 7.5, 9.0))
    path = tmp_path / "body.stl"
    write_stl_binary(path, tri)
    body = geometry.read_surface(path, S, 1.0, (-1.0, -1.0, -1.0))
    G, O = pair(size, 19, nu=0.05, dtype=dtype)
    G.init_equilibrium(1.0, (0.04, 0, 0))
    O.init_f(1.0, (0.04, 0, 0))
    bcs = [add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.04, 0, 0)),
           add_bc(G, O, _lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
           add_bc(G, O, _lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
           add_bc(G, O, _lib.BC_NOSLIP_MAP, mapv=body)]
    run(G, O, 40, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


def test_python_asl_layer_locomotive_style(tmp_path):
    """The Python mirror of ASL's names in a locomotive-style driver:
    readSurface(file, Block), LBGKTurbulence on D3Q15, a velocity map tunnel,
    a no-slip STL body, pressure-velocity ends and the surface force; against
    the oracle fed the same maps, bit for bit."""
    import asl_amd
    dx = 0.25
    grid = asl_amd.Block((40, 18, 22), dx, (-dx, -dx, -dx))
    path = tmp_path / "body.stl"
    # the body's shell stays cells away from the tunnel's: a velocity-map ghost
    # whose average would take in a body ghost is rejected (ASL would read that
    # ghost's stale populations)
    write_stl_binary(path, icosphere(0.6, (2.5, 1.4, 3.0)))
    body = asl_amd.readSurface(path, grid)
    size = tuple(s - 2 for s in grid.size)
    # tunnel as the drivers build it (not normalised): inside of a cylinder along x,
    # capped half a cell beyond the block's ends, cut by the ground plane y = 0
    pos = geometry.positions(size, dx, (0.0, 0.0, 0.0))
    r = (size[2] - 1) * dx / 2.1
    cx, cy, cz = 0.5 * (size[0] - 1) * dx, 0.25 * r, 0.5 * (size[2] - 1) * dx
    half = 0.5 * ((size[0] - 1) * dx + dx)
    tunnel = (r - np.sqrt((pos[1] - np.float32(cy)) ** 2 + (pos[2] - np.float32(cz)) ** 2)).astype(np.float32)
    tunnel = np.minimum(np.minimum(tunnel, np.float32(half) - np.abs(pos[0] - np.float32(cx))), pos[1])
    tunnel = tunnel.astype(np.float32)
    nu, u0 = 0.002, 0.1
    lb = asl_amd.LBGKTurbulence(asl_amd.Block(size), nu, asl_amd.d3q15())
    lb.init()
    asl_amd.LBGKUtilities(lb).initF((u0, 0.0, 0.0))
    bcs = [asl_amd.generateBCVelocity(lb, (u0, 0.0, 0.0), tunnel),
           asl_amd.generateBCNoSlip(lb, body),
           asl_amd.generateBCConstantPressureVelocity(lb, 1.0, (u0, 0.0, 0.0), ["X0", "XE"])]
    asl_amd.initAll(bcs)
    asl_amd.executeAll(bcs)
    force = asl_amd.generateComputeSurfaceForce(lb, body)
    force.init()
    for _ in range(60):
        lb.execute()
        asl_amd.executeAll(bcs)
    force.execute()
    O = oracle.OracleLattice(size, 15, nu=nu)
    O.set_turbulence(True)
    O.init_f(1.0, (u0, 0, 0))
    ids = [O.add_map_bc(O.VELOCITY_MAP, tunnel, u=(u0, 0, 0)), O.add_map_bc(O.NOSLIP_MAP, body),
           O.add_slice_bc(O.CONSTANT_PRESSURE_VELOCITY, _lib.X0 | _lib.XE, p=1.0, u=(u0, 0, 0))]
    for b in ids:
        O.execute_bc(b)
    O.run(60, ids)
    ct = lb.lattice.cell_types()
    comp = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    assert comp.sum() > 1000 and (body <= 0).sum() > 50
    assert np.array_equal(lb.getRho()[comp], O.rho()[comp].astype(np.float32))
    assert np.array_equal(lb.getVelocity()[:, comp], O.velocity()[:, comp].astype(np.float32))
    inner = (slice(None),) + (slice(1, -1),) * 3
    fo = O.surface_force(body)[inner]
    assert np.abs(fo).max() > 1e-3
    assert np.array_equal(force.field[inner], fo.astype(np.float32))
