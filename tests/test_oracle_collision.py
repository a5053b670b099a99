"""The oracle's collision vs textbook BGK, invariants and a closed-form decay.

ASL's kernel (src/num/aslLBGK.cxx:187-235) writes the equilibrium with the
momentum m = sum w_i c_i f_i of weight-normalised populations:
    feq_i = rho + 3 c_i.m + 4.5 (c_i.m)^2 / rho - 1.5 m.m / rho,
    f_i <- f_i (1 - w) + w feq_i,   w = 1 / (3 nu + 1/2).
The textbook form in velocities, rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u) with
u = m / rho, is algebraically equal; it is evaluated here independently in
numpy (FP64) and must agree to rounding.  A wrong sign, factor or dropped
term fails these tests, as does a wrong viscosity relation (decay test).
"""
import numpy as np
import pytest

from oracle import OracleLattice, template


def pull_periodic(f, c):
    """f_i(x) <- f_i(x - c_i) on the interior with periodic wrap (numpy)."""
    out = np.empty_like(f)
    for i in range(f.shape[0]):
        out[i] = np.roll(f[i], shift=tuple(c[i]), axis=(0, 1, 2))
    return out


def bgk_textbook(f, c, w_lat, nu, compressible=True):
    rho = np.einsum("i,i...->...", w_lat, f)
    m = np.einsum("i,ia,i...->a...", w_lat, c, f)
    om = 1.0 / (3.0 * nu + 0.5)
    out = np.empty_like(f)
    for i in range(f.shape[0]):
        if compressible:
            u = m / rho
            cu = np.einsum("a,a...->...", c[i], u)
            feq = rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(0))
        else:
            cm = np.einsum("a,a...->...", c[i], m)
            feq = rho + 3 * cm + 4.5 * cm * cm - 1.5 * (m * m).sum(0)
        out[i] = f[i] - om * (f[i] - feq)
    return out, rho, m


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("compressible", [True, False])
def test_one_step_matches_textbook_bgk(q, compressible):
    size = (6, 5, 7)
    c, w, _ = template(q)
    rng = np.random.default_rng(7)
    L = OracleLattice(size, q, fp64=True, nu=0.07, periodic=True)
    L.set_compressible(compressible)
    S = tuple(s + 2 for s in size)
    f0 = np.full((q,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + 0.05 * rng.standard_normal((q,) + size)
    L.write_f(f0)
    L.execute()
    fi = f0[:, 1:-1, 1:-1, 1:-1]
    ref, rho, m = bgk_textbook(pull_periodic(fi, c), c, w, 0.07, compressible)
    got = L.f()[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(got - ref).max() < 2e-15
    r = L.rho()[1:-1, 1:-1, 1:-1]
    v = L.velocity()[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(r - rho).max() < 2e-15
    assert np.abs(v - (m / rho if compressible else m)).max() < 2e-15


@pytest.mark.parametrize("q", [15, 19, 27])
def test_equilibrium_is_fixed_point(q):
    u0 = (0.03, -0.02, 0.01)
    L = OracleLattice((4, 5, 6), q, fp64=True, nu=0.05, periodic=True)
    L.init_f(1.02, u0)
    f0 = L.f()[:, 1:-1, 1:-1, 1:-1].copy()
    L.run(20)
    assert np.abs(L.f()[:, 1:-1, 1:-1, 1:-1] - f0).max() < 1e-14
    assert np.abs(L.rho()[1:-1, 1:-1, 1:-1] - 1.02).max() < 1e-14
    v = L.velocity()[:, 1:-1, 1:-1, 1:-1]
    for d in range(3):
        assert np.abs(v[d] - u0[d]).max() < 1e-15


@pytest.mark.parametrize("q", [15, 19, 27])
def test_equilibrium_moments(q):
    # sum w feq = rho, sum w c feq = rho u (closed form, exact up to rounding)
    size = (5, 4, 3)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(3)
    rho = 1 + 0.1 * rng.standard_normal(S)
    u = 0.05 * rng.standard_normal((3,) + S)
    L = OracleLattice(size, q, fp64=True)
    L.init_f(rho, u)
    c, w, _ = template(q)
    f = L.f()
    assert np.abs(np.einsum("i,i...->...", w, f) - rho).max() < 1e-15
    assert np.abs(np.einsum("i,ia,i...->a...", w, c, f) - rho * u).max() < 1e-15


@pytest.mark.parametrize("q", [15, 19, 27])
def test_mass_conserved_in_closed_box(q):
    size = (10, 9, 8)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(5)
    u = np.zeros((3,) + S)
    u[:, 1:-1, 1:-1, 1:-1] = 0.02 * rng.standard_normal((3,) + size)
    L = OracleLattice(size, q, fp64=True, nu=0.03)
    L.init_f(1.0, u)
    b = L.add_slice_bc(L.NOSLIP, 63)
    L.execute_bc(b)
    mask = np.zeros(S, dtype=np.uint8)
    mask[1:-1, 1:-1, 1:-1] = 1
    m0 = L.conservation(mask)[0]
    L.run(200, [b])
    m1 = L.conservation(mask)[0]
    assert abs(m1 - m0) / m0 < 1e-13


def test_momentum_conserved_periodic():
    size = (8, 7, 6)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(11)
    u = 0.03 * rng.standard_normal((3,) + S)
    L = OracleLattice(size, 19, fp64=True, nu=0.02, periodic=True)
    L.init_f(1.0, u)
    mask = np.zeros(S, dtype=np.uint8)
    mask[1:-1, 1:-1, 1:-1] = 1
    c0 = L.conservation(mask)
    L.run(100)
    c1 = L.conservation(mask)
    assert abs(c1[0] - c0[0]) < 1e-11
    assert np.abs(c1[1:4] - c0[1:4]).max() < 1e-12


@pytest.mark.parametrize("q", [19, 27])
def test_taylor_green_decay(q):
    """2D Taylor-Green vortex: kinetic energy decays as exp(-4 nu k^2 t)."""
    N, nu, u0, steps = 32, 0.05, 0.02, 300
    k = 2 * np.pi / N
    size = (N, N, 4)
    S = tuple(s + 2 for s in size)
    x = (np.arange(S[0]) - 1.0)[:, None, None]
    y = (np.arange(S[1]) - 1.0)[None, :, None]
    u = np.zeros((3,) + S)
    u[0] = u0 * np.sin(k * x) * np.cos(k * y) * np.ones(S)
    u[1] = -u0 * np.cos(k * x) * np.sin(k * y) * np.ones(S)
    p = -0.25 * u0 * u0 * (np.cos(2 * k * x) + np.cos(2 * k * y)) * np.ones(S)
    rho = 1.0 + 3.0 * p
    L = OracleLattice(size, q, fp64=True, nu=nu, periodic=True)
    L.init_f(rho, u)
    mask = np.zeros(S, dtype=np.uint8)
    mask[1:-1, 1:-1, 1:-1] = 1
    ke0 = L.conservation(mask)[4]
    L.run(steps)
    ke1 = L.conservation(mask)[4]
    expect = np.exp(-4 * nu * k * k * steps)
    # BGK at 32 cells per wavelength with an equilibrium start: ~1% faster decay
    assert ke1 / ke0 == pytest.approx(expect, rel=0.02)
    # a wrong viscosity relation (e.g. nu = (tau - 1/2)/6) would miss by far more
    assert abs(ke1 / ke0 - np.exp(-2 * nu * k * k * steps)) > 0.1


def rotate_sequential(m, om, t):
    """ASL's momentum += t * (omega x momentum), one component after the other."""
    m = m.copy()
    m[0] = m[0] + t * (om[1] * m[2] - om[2] * m[1])
    m[1] = m[1] + t * (om[2] * m[0] - om[0] * m[2])
    m[2] = m[2] + t * (om[0] * m[1] - om[1] * m[0])
    return m


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("compressible", [True, False])
def test_rotating_frame_one_step(q, compressible):
    """setOmega: textbook BGK with the momentum first moved by (2/w) omega x m.

    The expected value is built in numpy from the moments of f; the
    component-by-component update (ASL emits one statement per component) is
    distinguishable from the simultaneous one at these sizes, and so are a
    wrong sign, factor or operand order of the cross product.
    """
    size, nu, om = (5, 6, 4), 0.07, np.array([0.02, -0.03, 0.05])
    c, w, _ = template(q)
    rng = np.random.default_rng(5)
    L = OracleLattice(size, q, fp64=True, nu=nu, periodic=True)
    L.set_compressible(compressible)
    L.set_omega(om)
    S = tuple(s + 2 for s in size)
    f0 = np.full((q,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + 0.05 * rng.standard_normal((q,) + size)
    L.write_f(f0)
    L.execute()
    f = pull_periodic(f0[:, 1:-1, 1:-1, 1:-1], c)
    rho = np.einsum("i,i...->...", w, f)
    m = np.einsum("i,ia,i...->a...", w, c, f)
    wr = 1.0 / (3.0 * nu + 0.5)
    m2 = rotate_sequential(m, om, 2.0 / wr)
    ref = np.empty_like(f)
    for i in range(q):
        cm = np.einsum("a,a...->...", c[i], m2)
        if compressible:
            feq = rho + 3 * cm + 4.5 * cm * cm / rho - 1.5 * (m2 * m2).sum(0) / rho
        else:
            feq = rho + 3 * cm + 4.5 * cm * cm - 1.5 * (m2 * m2).sum(0)
        ref[i] = f[i] - wr * (f[i] - feq)
    got = L.f()[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(got - ref).max() < 4e-15
    v = L.velocity()[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(v - (m2 / rho if compressible else m2)).max() < 4e-15
    # the simultaneous update is a different (wrong) answer at this size
    m_sim = m + (2.0 / wr) * np.cross(om, m, axis=0)
    assert np.abs(m_sim - m2).max() > 1e-6
    # momentum after the collision: m + w (m' - m), i.e. m + 2 omega x m to first order
    mom = np.einsum("i,ia,i...->a...", w, c, got)
    assert np.abs(mom - (m + wr * (m2 - m))).max() < 4e-15


def test_rotating_frame_off_is_plain_bgk():
    size = (4, 4, 4)
    rng = np.random.default_rng(3)
    S = tuple(s + 2 for s in size)
    f0 = np.full((19,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + 0.05 * rng.standard_normal((19,) + size)
    out = []
    for om in (None, (0.0, 0.0, 0.0), "off"):
        L = OracleLattice(size, 19, fp64=True, nu=0.1, periodic=True)
        if om == "off":
            L.set_omega((0.1, 0.2, 0.3))
            L.set_omega(None)
        elif om is not None:
            L.set_omega(om)
        L.write_f(f0)
        L.execute()
        out.append(L.f()[:, 1:-1, 1:-1, 1:-1])
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def turbulence_rate(f, c, w_lat, nu, compressible):
    """LBGKTurbulence's per-cell relaxation rate (numpy, FP64), with fEq the
    textbook equilibrium of the cell's own moments (velocity form)."""
    rho = np.einsum("i,i...->...", w_lat, f)
    m = np.einsum("i,ia,i...->a...", w_lat, c, f)
    d2 = np.zeros_like(rho)
    for i in range(f.shape[0]):
        if compressible:
            u = m / rho
            cu = np.einsum("a,a...->...", c[i], u)
            feq = rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(0))
        else:
            cm = np.einsum("a,a...->...", c[i], m)
            feq = rho + 3 * cm + 4.5 * cm * cm - 1.5 * (m * m).sum(0)
        d2 += (f[i] - feq) ** 2
    r = 4.0 * d2 / (f * f).sum(0)
    plain = 1.0 / (3.0 * nu + 0.5)
    return np.where(r > 1.0, 1.0 / (3.0 * np.minimum(nu * r, 0.5) + 0.5), plain), r


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("compressible", [True, False])
def test_turbulence_one_step(q, compressible):
    """LBGKTurbulence (src/num/aslLBGK.cxx:298-355): one step against a numpy
    model that applies the textbook BGK with the per-cell rate
    w = 1 / (3 min(nu r, 1/2) + 1/2) where r = 4 |f - feq|^2 / |f|^2 > 1.
    The per-cell amplitudes put cells in all three regimes (plain rate,
    limited rate, clamped at 1/2); a dropped factor 4, a swapped select or a
    missing clamp moves many cells by far more than the tolerance."""
    size, nu = (7, 6, 5), 0.2
    c, w, _ = template(q)
    rng = np.random.default_rng(40 + q)
    L = OracleLattice(size, q, fp64=True, nu=nu, periodic=True)
    L.set_compressible(compressible)
    L.set_turbulence(True)
    S = tuple(s + 2 for s in size)
    amp = rng.uniform(0.01, 1.6, size)
    f0 = np.full((q,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + amp * rng.standard_normal((q,) + size)
    L.write_f(f0)
    L.execute()
    f = pull_periodic(f0[:, 1:-1, 1:-1, 1:-1], c)
    wc, r = turbulence_rate(f, c, w, nu, compressible)
    assert (r <= 1).sum() > 10 and ((r > 1) & (nu * r < 0.5)).sum() > 10 and (nu * r >= 0.5).sum() > 10
    rho = np.einsum("i,i...->...", w, f)
    m = np.einsum("i,ia,i...->a...", w, c, f)
    ref = np.empty_like(f)
    for i in range(q):
        cm = np.einsum("a,a...->...", c[i], m)
        if compressible:
            feq = rho + 3 * cm + 4.5 * cm * cm / rho - 1.5 * (m * m).sum(0) / rho
        else:
            feq = rho + 3 * cm + 4.5 * cm * cm - 1.5 * (m * m).sum(0)
        ref[i] = f[i] * (1 - wc) + wc * feq
    got = L.f()[:, 1:-1, 1:-1, 1:-1]
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() < 1e-13 * scale


@pytest.mark.parametrize("q", [15, 19, 27])
def test_turbulence_near_equilibrium_is_plain_lbgk(q):
    """Where every cell has r <= 1 the limiter keeps 1 / (3 nu + 1/2): the run
    is bit-identical to plain LBGK (FP32)."""
    size = (6, 5, 7)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(8)
    u = 0.03 * rng.standard_normal((3,) + S)
    runs = []
    for turb in (False, True):
        L = OracleLattice(size, q, fp64=False, nu=0.05, periodic=True)
        L.set_turbulence(turb)
        L.init_f(1.0, u)
        L.run(30)
        runs.append(L.f())
    assert np.array_equal(runs[0], runs[1])


def test_turbulence_limits_the_rate_far_from_equilibrium():
    """A cell far from equilibrium relaxes with w = 1/(3 * 0.5 + 0.5) = 1/2 when
    nu r >= 1/2, whatever nu is: two viscosities give the same step there."""
    size = (4, 4, 4)
    S = tuple(s + 2 for s in size)
    q = 19
    rng = np.random.default_rng(2)
    f0 = np.full((q,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + rng.uniform(0.01, 3.0, size) * rng.standard_normal((q,) + size)
    out = []
    for nu in (0.3, 0.6):
        L = OracleLattice(size, q, fp64=True, nu=nu, periodic=True)
        L.set_turbulence(True)
        L.write_f(f0)
        L.execute()
        out.append(L.f()[:, 1:-1, 1:-1, 1:-1])
    c, w, _ = template(q)
    f = pull_periodic(f0[:, 1:-1, 1:-1, 1:-1], c)
    _, r = turbulence_rate(f, c, w, 0.3, True)
    clamped = 0.3 * r >= 0.5
    assert clamped.sum() > 5
    np.testing.assert_array_equal(out[0][:, clamped], out[1][:, clamped])
    assert not np.array_equal(out[0][:, ~clamped], out[1][:, ~clamped])
