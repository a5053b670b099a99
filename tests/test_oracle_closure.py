"""The closure's CPU restatements: network, training step, features, collision.

The network has no ASL counterpart; these tests pin the oracle by what
mathematics fixes: finite differences for the gradient, Adam's first step,
the moments of an equilibrium (no non-equilibrium part), an independent
numpy evaluation of the whole cell update, and the reductions to plain BGK
(zero correction, constant correction).
"""
import numpy as np
import pytest

from oracle import OracleLattice, mlp, template


def test_gradient_matches_finite_differences():
    p = mlp.init_params(3).astype(np.float64)
    rng = np.random.default_rng(1)
    x, t = rng.standard_normal((40, 7)), rng.standard_normal(40)
    _, g = mlp.loss_and_grad(p, x, t)
    for i in range(0, mlp.NPARAM, 37):
        e = np.zeros_like(p)
        e[i] = 1e-6
        num = (mlp.loss_and_grad(p + e, x, t)[0] - mlp.loss_and_grad(p - e, x, t)[0]) / 2e-6
        assert abs(num - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), i


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 2.0])
    p1, m, v = mlp.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p1, p - 0.01 * np.sign(g), rtol=0, atol=1e-7)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-40], dtype=np.float32)
    got = mlp.bf16_round(x)
    # ties go to even; a subnormal keeps its top 16 bits (0x000116C2 -> 0x00010000)
    expect = np.array([1.0, 1.0, 1.0078125, -3.140625, 65536.0, 0.0], dtype=np.float32)
    expect[5] = np.array([0x00010000], dtype=np.uint32).view(np.float32)[0]
    np.testing.assert_array_equal(got, expect)


def numpy_features(f, c, w):
    rho = np.einsum("i,i...->...", w, f)
    m = np.einsum("i,ia,i...->a...", w, c, f)
    P = np.einsum("i,ia,ib,i...->ab...", w, c, c, f)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    x = [P[a, b] - (rho / 3.0 if a == b else 0.0) - m[a] * m[b] / rho for a, b in pairs]
    return np.stack(x + [rho - 1.0], axis=-1), rho, m


def test_features_of_equilibrium_vanish():
    L = OracleLattice((6, 5, 4), 19, fp64=False, nu=0.05, periodic=True)
    L.init_f(1.03, (0.04, -0.02, 0.01))
    x = L.closure_features()
    assert np.abs(x[..., :6]).max() < 2e-7
    np.testing.assert_allclose(x[..., 6], 0.03, atol=2e-7)


def test_features_match_numpy():
    c, w, _ = template(19)
    size = (5, 6, 4)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(2)
    L = OracleLattice(size, 19, fp64=False, nu=0.05, periodic=True)
    f0 = np.full((19,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = 1.0 + 0.05 * rng.standard_normal((19,) + size)
    L.write_f(f0)
    x = L.closure_features()
    fi = f0[:, 1:-1, 1:-1, 1:-1].astype(np.float32).astype(np.float64)
    # the next collision of a periodic lattice pulls f_i from x - c_i
    pulled = np.stack([np.roll(fi[i], tuple(c[i]), axis=(0, 1, 2)) for i in range(19)])
    ref, _, _ = numpy_features(pulled, c, w)
    assert np.abs(x - ref).max() < 2e-6


def closure_step_numpy(f, c, w, nu, params, mu, sc, out):
    """One periodic step with the closure, FP64 numpy (BF16-rounded parameters)."""
    pulled = np.stack([np.roll(f[i], tuple(c[i]), axis=(0, 1, 2)) for i in range(19)])
    x, rho, m = numpy_features(pulled, c, w)
    xn = (x - mu) * sc
    y, _ = mlp.forward(mlp.bf16_round(params).astype(np.float64), xn.reshape(-1, 7))
    tau0 = 3.0 * nu + 0.5
    tau = np.maximum(tau0 + y.reshape(rho.shape) * out, 0.5 + 0.5 * (tau0 - 0.5))
    wr = 1.0 / tau
    outf = np.empty_like(pulled)
    for i in range(19):
        cm = np.einsum("a,a...->...", c[i], m)
        feq = rho + 3 * cm + 4.5 * cm * cm / rho - 1.5 * (m * m).sum(0) / rho
        outf[i] = pulled[i] - wr * (pulled[i] - feq)
    return outf, tau


@pytest.fixture
def random_state():
    size = (6, 5, 7)
    S = tuple(s + 2 for s in size)
    rng = np.random.default_rng(4)
    f0 = np.full((19,) + S, np.nan)
    f0[:, 1:-1, 1:-1, 1:-1] = (1.0 + 0.02 * rng.standard_normal((19,) + size)).astype(np.float32)
    return size, f0


def run_oracle(size, f0, nu, closure=None, mode=0):
    L = OracleLattice(size, 19, fp64=False, nu=nu, periodic=True)
    if closure is not None:
        L.set_closure(*closure)
        L.set_closure_mode(mode)
    L.write_f(f0)
    L.execute()
    return L.f()[:, 1:-1, 1:-1, 1:-1]


def test_closure_step_matches_numpy(random_state):
    size, f0 = random_state
    c, w, _ = template(19)
    nu = 0.01
    p = mlp.init_params(7)
    p[mlp.NPARAM - 1] = 0.3  # b3: most cells get a correction
    mu = np.zeros(7, np.float32)
    sc = np.full(7, 40.0, np.float32)
    got = run_oracle(size, f0, nu, (p, mu, sc, 0.05))
    ref, tau = closure_step_numpy(f0[:, 1:-1, 1:-1, 1:-1].astype(np.float64), c, w, nu, p, mu, sc, 0.05)
    assert (tau > 3 * nu + 0.5 + 1e-4).mean() > 0.3  # the network does act
    assert np.abs(got - ref).max() < 3e-6


def bf16_model_tau(x, params, mu, sc, out, tau0):
    """The matrix-core evaluation in numpy: BF16 inputs and first hidden layer, exact sums."""
    W1, b1, W2, b2, w3, b3 = mlp.unpack(mlp.bf16_round(params).astype(np.float64))
    xb = mlp.bf16_round(((x.astype(np.float32) - mu) * sc).astype(np.float32)).astype(np.float64)
    h1 = mlp.bf16_round(np.maximum(xb @ W1.T + b1, 0).astype(np.float32)).astype(np.float64)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ w3 + b3
    return np.maximum(tau0 + y * out, 0.5 + 0.5 * (tau0 - 0.5))


def test_bf16_evaluation_matches_numpy_model(random_state):
    size, f0 = random_state
    c, w, _ = template(19)
    nu = 0.01
    p = mlp.init_params(7)
    p[mlp.NPARAM - 1] = 0.3
    mu, sc, out = np.zeros(7, np.float32), np.full(7, 40.0, np.float32), 0.05
    got = run_oracle(size, f0, nu, (p, mu, sc, out), mode=1)
    fp32 = run_oracle(size, f0, nu, (p, mu, sc, out), mode=0)
    fi = f0[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    pulled = np.stack([np.roll(fi[i], tuple(c[i]), axis=(0, 1, 2)) for i in range(19)])
    x, rho, m = numpy_features(pulled, c, w)
    tau = bf16_model_tau(x.reshape(-1, 7), p, mu, sc, out, 3 * nu + 0.5).reshape(rho.shape)
    ref = np.empty_like(pulled)
    for i in range(19):
        cm = np.einsum("a,a...->...", c[i], m)
        feq = rho + 3 * cm + 4.5 * cm * cm / rho - 1.5 * (m * m).sum(0) / rho
        ref[i] = pulled[i] - (pulled[i] - feq) / tau
    assert np.abs(got - ref).max() < 3e-6
    assert np.abs(got - fp32).max() > 1e-8  # the two evaluations do differ


def test_zero_correction_is_plain_bgk(random_state):
    size, f0 = random_state
    p = np.zeros(mlp.NPARAM, np.float32)  # y = 0 everywhere: no correction
    plain = run_oracle(size, f0, 0.02)
    cl = run_oracle(size, f0, 0.02, (p, np.zeros(7), np.ones(7), 0.1))
    assert np.array_equal(plain, cl)


def test_negative_correction_is_floored_at_half_the_viscosity(random_state):
    size, f0 = random_state
    p = np.zeros(mlp.NPARAM, np.float32)
    p[mlp.NPARAM - 1] = -1.0  # tau0 - 1 is below the floor 1/2 + (tau0 - 1/2) / 2
    cl = run_oracle(size, f0, 0.02, (p, np.zeros(7), np.ones(7), 1.0))
    plain = run_oracle(size, f0, 0.01)
    assert np.abs(cl - plain).max() < 1e-6


def test_constant_correction_is_bgk_with_larger_tau(random_state):
    size, f0 = random_state
    p = np.zeros(mlp.NPARAM, np.float32)
    p[mlp.NPARAM - 1] = 0.5  # y = 0.5: tau = tau0 + 0.5 * 0.2 = tau0 + 0.1
    cl = run_oracle(size, f0, 0.02, (p, np.zeros(7), np.ones(7), 0.2))
    plain = run_oracle(size, f0, 0.02 + 0.1 / 3.0)
    assert np.abs(cl - plain).max() < 1e-6
