"""Host-side data of the closure (asl_amd.closure): initial field, pressure, filters."""
import numpy as np

from asl_amd import closure


def test_turbulence_field_statistics():
    u = closure.turbulence_field(32, seed=4, u_rms=0.03)
    assert u.shape == (3, 32, 32, 32) and u.dtype == np.float32
    np.testing.assert_allclose(np.sqrt((u.astype(np.float64) ** 2).sum(0).mean() / 3.0), 0.03, rtol=1e-5)
    assert np.abs(u.mean(axis=(1, 2, 3))).max() < 1e-8
    # divergence-free in spectral space
    k = np.fft.fftfreq(32, d=1.0 / 32)
    K = np.meshgrid(k, k, k, indexing="ij")
    div = sum(K[i] * np.fft.fftn(u[i].astype(np.float64)) for i in range(3))
    assert np.abs(div).max() < 1e-5 * np.abs(np.fft.fftn(u[0].astype(np.float64))).max() * 32
    # same seed, same field; another seed, another field
    assert np.array_equal(u, closure.turbulence_field(32, seed=4, u_rms=0.03))
    assert not np.array_equal(u, closure.turbulence_field(32, seed=5, u_rms=0.03))


def test_pressure_of_taylor_green_vortex():
    """u = sin x cos y, v = -cos x sin y has p = (cos 2x + cos 2y) / 4."""
    n = 32
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0 * X]).astype(np.float32)
    p = closure.pressure_field(u)
    assert np.abs(p - (np.cos(2 * X) + np.cos(2 * Y)) / 4).max() < 1e-6


def test_box_filter_and_ghosts():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 8, 12, 16))
    b = closure.box_filter(a, 4)
    assert b.shape == (2, 2, 3, 4)
    np.testing.assert_allclose(b.mean(axis=(1, 2, 3)), a.mean(axis=(1, 2, 3)))
    np.testing.assert_allclose(b[1, 1, 2, 3], a[1, 4:8, 8:12, 12:16].mean())
    g = closure.with_ghosts(a)
    assert g.shape == (2, 10, 14, 18)
    assert np.array_equal(g[:, 0, 1:-1, 1:-1], a[:, -1]) and np.array_equal(g[:, -1, 1:-1, 1:-1], a[:, 0])


def test_eddy_viscosity_of_a_resolved_field_is_small():
    """A field with all its energy at scales far above the filter has almost no subgrid stress."""
    n = 64
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = 0.03 * np.stack([np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0 * X])
    nut, ub = closure.eddy_viscosity(u, 4)
    assert ub.shape == (3, 16, 16, 16)
    assert np.abs(nut).max() < 1e-3  # coarse lattice units; the filter width is 4 fine cells
