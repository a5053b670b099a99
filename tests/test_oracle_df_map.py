"""The oracle's distance-function maps (initData(map, DF), src/aslGeomInc.cxx:
187-221) against closed forms, CSG identities and the numpy builders of
asl_amd.geometry (a separate evaluation of the same ASL formulas,
src/math/aslDistanceFunction.cxx:55-164).  Integer geometry makes squares and
square roots of perfect squares exact, so most checks are bit for bit."""
import numpy as np
import pytest

import oracle
from asl_amd import geometry as g

F = np.float32


def grid_pos(shape, pos, dx, dtype=F):
    return [dtype(pos[d]) + dtype(dx) * np.arange(shape[d], dtype=dtype).reshape(
        [-1 if k == d else 1 for k in range(3)]) for d in range(3)]


@pytest.mark.parametrize("fp64", [False, True])
def test_sphere_exact_on_integer_geometry(fp64):
    shape, c, r = (11, 9, 13), (5.0, 4.0, 6.0), 3.0
    got = oracle.df_map(g.df_sphere(r, c), shape, (0, 0, 0), 1.0, fp64)
    I = np.indices(shape).astype(np.float64)
    d2 = sum((I[k] - c[k]) ** 2 for k in range(3))
    ref = np.sqrt(d2.astype(np.float64 if fp64 else F)) - (np.float64(r) if fp64 else F(r))
    assert np.array_equal(got, ref)
    # on the sphere (d2 = 9): exactly zero
    assert (got[d2 == 9] == 0).all() and (d2 == 9).sum() > 10


def test_matches_numpy_builders_of_the_product():
    """geometry.sphere / cylinder_inf / normalize / CSG (numpy, float) against the
    oracle's program evaluation at the same positions, bit for bit."""
    size, dx, position = (14, 12, 10), 0.5, (0.25, -1.0, 0.5)
    S = tuple(s + 2 for s in size)
    p0 = [position[d] - dx for d in range(3)]
    sph = g.normalize(g.sphere(size, 1.7, (3.1, 2.2, 2.9), dx, position), dx)
    cyl = g.cylinder_inf(size, 1.2, (0.3, 1.0, -0.4), (2.0, 1.5, 3.0), dx, position)
    ref = g.union(sph, g.complement(cyl))
    prog = g.df_or(g.df_normalize(g.df_sphere(1.7, (3.1, 2.2, 2.9)), dx),
                   g.df_neg(g.df_cylinder_inf(1.2, (0.3, 1.0, -0.4), (2.0, 1.5, 3.0))))
    got = oracle.df_map(prog, S, p0, dx)
    assert np.array_equal(got, ref)


def test_plane_and_csg_identities():
    shape = (9, 8, 7)
    A = g.df_plane((0.0, 2.0, 0.0), (0.0, 3.0, 0.0))       # y - 3
    B = g.df_sphere(2.5, (4.0, 4.0, 3.0))
    a, b = oracle.df_map(A, shape, (0, 0, 0), 1.0), oracle.df_map(B, shape, (0, 0, 0), 1.0)
    I = np.indices(shape)
    assert np.array_equal(a, (I[1] - 3).astype(F))
    assert np.array_equal(oracle.df_map(g.df_and(A, B), shape, (0, 0, 0), 1.0), np.maximum(a, b))
    assert np.array_equal(oracle.df_map(g.df_or(A, B), shape, (0, 0, 0), 1.0), np.minimum(a, b))
    assert np.array_equal(oracle.df_map(g.df_neg(A), shape, (0, 0, 0), 1.0), -a)
    n = oracle.df_map(g.df_normalize(B, 0.5), shape, (0, 0, 0), 1.0)
    assert np.array_equal(n, np.clip(F(1 / 0.5 / 1.8) * b, -1, 1))


def test_finite_cylinder_and_cone_closed_forms():
    shape = (16, 16, 16)
    # cylinder of radius 3 along x through (8, 8, 8), length 8: caps at x = 4, 12
    cyl = oracle.df_map(g.df_cylinder(3.0, (8.0, 0.0, 0.0), (8.0, 8.0, 8.0)), shape, (0, 0, 0), 1.0)
    I = np.indices(shape).astype(np.float64)
    rad = np.sqrt((I[1] - 8) ** 2 + (I[2] - 8) ** 2) - 3
    cap = np.maximum(I[0] - 12, 4 - I[0])
    np.testing.assert_allclose(cyl, np.maximum(rad, cap), atol=2e-6)
    # cone with apex (2, 8, 8), axis +x, length 10, base radius 5: tan(th) = tan(0.5)
    cone = oracle.df_map(g.df_cone(5.0, (10.0, 0.0, 0.0), (2.0, 8.0, 8.0)), shape, (0, 0, 0), 1.0)
    proj = I[0] - 2
    body = np.sqrt((I[1] - 8) ** 2 + (I[2] - 8) ** 2) - np.tan(0.5) * proj
    ref = np.maximum(np.maximum(body, proj - 10), -proj)
    np.testing.assert_allclose(cone, ref, atol=4e-6)
    assert (cone < 0).sum() > 50


def test_grid_positions_and_spacing():
    prog = g.df_plane((1.0, 0.0, 0.0), (0.0, 0.0, 0.0))  # x
    got = oracle.df_map(prog, (5, 2, 3), (-0.75, 0, 0), 0.25)
    assert np.array_equal(got[:, 0, 0], F(-0.75) + F(0.25) * np.arange(5, dtype=F))
