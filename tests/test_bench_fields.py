"""The C2 workload's counter-based perturbation field (bench.py)."""
import numpy as np

import bench


def test_splitmix64_known_answers():
    # splitmix64 seeded with 0: first outputs 0xE220A8397B1DCDAF,
    # 0x6E789E6AA1B965F4, 0x06C45D188009454F (Vigna's reference generator)
    top = np.array([0xE220A8, 0x6E789E, 0x06C45D], dtype=np.float64)
    expect = 2.0 * top / 2.0 ** 24 - 1.0
    got = bench.splitmix64_uniform(0, 3)
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got.astype(np.float64), expect)


def test_perturbation_range_and_mean():
    x = bench.splitmix64_uniform(20261202, 1 << 20).astype(np.float64)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert abs(x.mean()) < 5e-3 and abs(x.var() - 1.0 / 3.0) < 5e-3


def test_setup_fields_shape_and_solid():
    mapv, u = bench.setup_fields()
    S = tuple(s + 2 for s in bench.SIZE)
    assert mapv.shape == S and u.shape == (3,) + S
    assert np.all(u[:, mapv <= 0] == 0)
    fluid = mapv > 0
    assert np.all(np.abs(u[0][fluid] - bench.U_IN) <= 1.001e-3)
    assert np.all(np.abs(u[1][fluid]) <= 1.001e-3)
