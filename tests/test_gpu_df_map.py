"""Distance-function maps on the GPU (aslhip_df_map) against the oracle, bit
for bit: random CSG trees of spheres, cylinders, cones and planes with
normalisation, in float and double, on shifted grids."""
import numpy as np
import pytest

import asl_amd
import oracle
from asl_amd import geometry as g

pytestmark = pytest.mark.gpu


def random_tree(rng, depth):
    if depth == 0 or rng.random() < 0.25:
        k = rng.integers(5)
        c = rng.uniform(0, 10, 3)
        if k == 0:
            return g.df_sphere(rng.uniform(1, 4), c)
        if k == 1:
            return g.df_cylinder_inf(rng.uniform(0.5, 3), rng.normal(size=3), c)
        if k == 2:
            return g.df_cylinder(rng.uniform(0.5, 3), rng.normal(size=3) * 4, c)
        if k == 3:
            return g.df_cone(rng.uniform(1, 3), rng.normal(size=3) * 5, c)
        return g.df_plane(rng.normal(size=3), c)
    op = rng.integers(4)
    if op == 0:
        return g.df_and(random_tree(rng, depth - 1), random_tree(rng, depth - 1))
    if op == 1:
        return g.df_or(random_tree(rng, depth - 1), random_tree(rng, depth - 1))
    if op == 2:
        return g.df_neg(random_tree(rng, depth - 1))
    return g.df_normalize(random_tree(rng, depth - 1), rng.uniform(0.2, 1.0))


@pytest.mark.parametrize("fp64", [False, True])
def test_random_trees_match_oracle(fp64):
    rng = np.random.default_rng(5 + fp64)
    for k in range(12):
        prog = random_tree(rng, 3)
        shape = tuple(int(s) for s in rng.integers(6, 30, 3))
        pos, dx = rng.uniform(-2, 2, 3), rng.uniform(0.2, 0.7)
        got = g.df_map(prog, shape, pos, dx, fp64)
        ref = oracle.df_map(prog, shape, pos, dx, fp64)
        assert np.array_equal(got, ref), (k, prog)


def test_the_benchmark_sphere():
    """The 256x128x128 channel's obstacle: normalize(sphere(16, (64,64,64)), 1) on
    the ghost-inclusive grid, equal to the numpy builder used by the benchmark."""
    size = (256, 128, 128)
    S = tuple(s + 2 for s in size)
    got = g.df_map(g.df_normalize(g.df_sphere(16.0, (64.0, 64.0, 64.0)), 1.0), S, (-1.0, -1.0, -1.0), 1.0)
    ref = g.normalize(g.sphere(size, 16.0, (64.0, 64.0, 64.0)), 1.0)
    assert np.array_equal(got, ref)


def test_malformed_programs_raise():
    for prog in ([5.0], [1.0, 0, 0, 0], [1.0, 0, 0, 0, 1, 1.0, 0, 0, 0, 1], [9.0], [1.0, 0, 0, 0, np.nan]):
        with pytest.raises(asl_amd.ASLError):
            g.df_map(prog, (4, 4, 4), (0, 0, 0), 1.0)
