"""The oracle's visualisation fills (generateBCNoSlipVel / NoSlipRho /
VelocityVel) against closed forms on simple geometry.

* a uniform stream is mirrored: every boundary node with a fluid neighbour
  gets -U (the no-slip midpoint), one without gets 0;
* a flat wall: the fill is the laplace-weighted (velocity) or
  normal-weighted (density) mean over the fluid neighbours, whose weights
  are written out by hand for D3Q15 and D3Q19 (a wrong normal sign, a wrong
  axis or the wrong stencil gives other values);
* a diagonal wall: the normal of a plane is fixed by geometry, so the
  density weights (n . c_i/|c_i| + 1)^2 follow from it;
* a position function: constant or axis x (position - centre) at the
  physical node positions of the block.
"""
import numpy as np
import pytest

import oracle

SIZE = (6, 7, 8)


def lattice(q, rho, u):
    O = oracle.OracleLattice(SIZE, q, fp64=True, nu=0.1)
    if np.ndim(rho) and np.ndim(u) == 1:
        u = np.broadcast_to(np.asarray(u, dtype=np.float64).reshape(3, 1, 1, 1), (3,) + np.shape(rho))
    O.init_f(rho, u)
    return O


def idx_grid():
    S = tuple(s + 2 for s in SIZE)
    return np.meshgrid(*[np.arange(s, dtype=np.float64) for s in S], indexing="ij")


def sphere_map():
    X, Y, Z = idx_grid()
    return (np.sqrt((X - 3.5) ** 2 + (Y - 4.0) ** 2 + (Z - 4.5) ** 2) - 1.7).astype(np.float32)


def guard(m):
    return (m <= 0) & (m > -0.9999)


@pytest.mark.parametrize("q", [15, 19])
def test_noslip_velocity_mirrors_a_uniform_stream(q):
    U = (0.03, -0.02, 0.01)
    O = lattice(q, 1.0, U)
    m = sphere_map()
    O.vis_fill(1, m)
    v = O.velocity()
    c, _, _ = oracle.template(q)
    S = m.shape
    g = guard(m)
    has_fluid = np.zeros(S, dtype=bool)
    for ci in c[1:]:
        sh = np.roll(m, tuple(-ci), axis=(0, 1, 2))  # the sphere is far from the block edges
        has_fluid |= sh > 0
    for d in range(3):
        assert np.allclose(v[d][g & has_fluid], -U[d], rtol=1e-14, atol=0)
        assert np.all(v[d][g & ~has_fluid] == 0.0)
        assert np.all(v[d][~g] == U[d])
    assert (g & has_fluid).sum() >= 16


def wall_setup(q, field):
    """Map = I_x - 1.5: the layer I_x = 1 is the guarded ghost layer, I_x >= 2 fluid."""
    X, Y, Z = idx_grid()
    m = (X - 1.5).astype(np.float32)
    f = field(Y, Z)
    return m, f


def tangential(Y, Z):
    return 1.0 + 0.01 * np.sin(1.3 * Y) + 0.02 * np.cos(0.7 * Z) + 0.003 * Y * Z


def hand_wall_mean(f, q, w_axis, w_off):
    """Weighted mean over the fluid neighbours (I_x = 2) of the guard cell
    (1, y, z): the axis neighbour (2, y, z) and, for D3Q19, the four edge
    neighbours (2, y +- 1, z), (2, y, z +- 1); for D3Q15 the four corners
    (2, y +- 1, z +- 1)."""
    S = f.shape
    out = np.full(S[1:], np.nan)
    for y in range(1, S[1] - 1):
        for z in range(1, S[2] - 1):
            if q == 19:
                offs = [(1, 0), (-1, 0), (0, 1), (0, -1)]
            else:
                offs = [(1, 1), (1, -1), (-1, 1), (-1, -1)]
            num = w_axis * f[2, y, z] + sum(w_off * f[2, y + a, z + b] for a, b in offs)
            out[y, z] = num / (w_axis + 4 * w_off)
    return out


@pytest.mark.parametrize("q", [15, 19])
def test_noslip_velocity_flat_wall(q):
    m, f = wall_setup(q, tangential)
    u = np.zeros((3,) + m.shape)
    u[1] = f - 1.0
    O = lattice(q, 1.0, u)
    O.vis_fill(1, m)
    v = O.velocity()
    # laplasCoefs: D3Q19 axis 1/3, edges 1/6; D3Q15 axis 2/3, corners 1/12
    la, lo = (1.0 / 3.0, 1.0 / 6.0) if q == 19 else (2.0 / 3.0, 1.0 / 12.0)
    want = -hand_wall_mean(u[1], q, la, lo)
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(v[1][1][inner], want[inner], rtol=1e-13, atol=1e-16)
    assert np.all(v[0][1][inner] == 0.0)


@pytest.mark.parametrize("q", [15, 19])
def test_noslip_density_flat_wall(q):
    m, f = wall_setup(q, tangential)
    O = lattice(q, f, (0.0, 0.0, 0.0))
    O.vis_fill(2, m)
    r = O.rho()
    # the map normal is +x: weight (1 + 1)^2 = 4 on the axis neighbour,
    # (1 + 1/sqrt(2))^2 on D3Q19 edges, (1 + 1/sqrt(3))^2 on D3Q15 corners
    wo = (1.0 + 1.0 / np.sqrt(2.0)) ** 2 if q == 19 else (1.0 + 1.0 / np.sqrt(3.0)) ** 2
    want = hand_wall_mean(f, q, 4.0, wo)
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(r[1][inner], want[inner], rtol=1e-13, atol=0)
    # a wall whose normal pointed the other way would weight these neighbours
    # with (1 - 1)^2 = 0 and (1 - 1/sqrt(2))^2: the fill differs
    wrong = hand_wall_mean(f, q, 0.0, (1.0 - 1.0 / np.sqrt(2.0 if q == 19 else 3.0)) ** 2)
    assert np.abs(r[1][inner] - wrong[inner]).max() > 1e-4


def test_noslip_density_diagonal_wall():
    q = 19
    X, Y, Z = idx_grid()
    m = ((X + Y) / np.sqrt(2.0) - 3.3).astype(np.float32)
    f = 1.0 + 0.05 * np.sin(0.9 * X) * np.cos(0.4 * Y) + 0.01 * Z
    O = lattice(q, f, (0.0, 0.0, 0.0))
    O.vis_fill(2, m)
    r = O.rho()
    c, _, _ = oracle.template(q)
    n = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)  # the plane's normal, towards the fluid
    S = m.shape
    checked = 0
    for x in range(1, S[0] - 1):
        for y in range(1, S[1] - 1):
            for z in range(1, S[2] - 1):
                if not (-0.9999 < m[x, y, z] <= 0):
                    continue
                num = den = 0.0
                for ci in c[1:]:
                    nb = (x + ci[0], y + ci[1], z + ci[2])
                    if m[nb] > 0:
                        wgt = (n @ (ci / np.linalg.norm(ci)) + 1.0) ** 2
                        num += wgt * f[nb]
                        den += wgt
                if den > 0.1:
                    assert abs(r[x, y, z] - num / den) <= 1e-12 * abs(num / den)
                    checked += 1
    assert checked >= 20


def test_velocity_position_function():
    q = 19
    O = lattice(q, 1.0, (0.0, 0.0, 0.0))
    dx, pos = 0.5, (1.0, -2.0, 0.25)
    O.set_geometry(dx, pos)
    m = sphere_map()
    axis, centre = np.array([0.1, -0.3, 0.2]), np.array([2.0, 0.0, 1.5])
    O.vis_fill(3, m, rot=np.concatenate([axis, centre]))
    v = O.velocity()
    X, Y, Z = idx_grid()
    P = np.stack([pos[0] - dx + dx * X, pos[1] - dx + dx * Y, pos[2] - dx + dx * Z])
    want = np.cross(axis, np.moveaxis(P, 0, -1) - centre)
    g = guard(m)
    for d in range(3):
        assert np.allclose(v[d][g], want[..., d][g], rtol=1e-13, atol=1e-15)
        assert np.all(v[d][~g] == 0.0)
    U = (0.01, 0.02, -0.04)
    O2 = lattice(q, 1.0, (0.0, 0.0, 0.0))
    O2.vis_fill(3, m, u=U)
    v2 = O2.velocity()
    for d in range(3):
        assert np.all(v2[d][g] == U[d])


def test_d3q27_has_no_stencil_coefficients():
    O = oracle.OracleLattice(SIZE, 27, fp64=True, nu=0.1)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    m = sphere_map()
    for kind in (1, 2):
        with pytest.raises(ValueError):
            O.vis_fill(kind, m)
    O.vis_fill(3, m, u=(0.1, 0.0, 0.0))
