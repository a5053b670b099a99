"""The learned closure on the GPU: features, fused collision and training kernels.

* features and the closure step are compared with the CPU oracle bit for bit
  (same operation order, FP32);
* the training kernels' gradient and Adam steps are compared with the FP64
  numpy restatement (oracle/mlp.py) at FP32 tolerance;
* a short training run on box-filtered fine-lattice data lowers the loss.
"""
import numpy as np
import pytest

import asl_amd
import oracle
from asl_amd import _lib, closure
from oracle import mlp

pytestmark = pytest.mark.gpu
WALLS = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0


def closure_args(seed=7):
    p = mlp.init_params(seed)
    p[mlp.NPARAM - 1] = 0.3
    return p, np.zeros(7, np.float32), np.full(7, 40.0, np.float32), 0.05


def periodic_pair(n=(10, 9, 8), dtype="f32"):
    rng = np.random.default_rng(3)
    S = tuple(s + 2 for s in n)
    u = 0.03 * rng.standard_normal((3,) + S)
    G = asl_amd.Lattice(n, 19, dtype=dtype, periodic=True, nu=0.01)
    O = oracle.OracleLattice(n, 19, fp64=False, nu=0.01, periodic=True)
    G.init_equilibrium(1.0, u.astype(np.float32))
    O.init_f(1.0, u.astype(np.float32).astype(np.float64))
    return G, O


def test_features_bitwise():
    G, O = periodic_pair()
    for steps in (0, 3, 4):  # both parities of the in-place layout
        G.step(steps)
        O.run(steps)
        assert np.array_equal(G.closure_features(), O.closure_features())


def test_closure_step_periodic_bitwise():
    G, O = periodic_pair()
    args = closure_args()
    G.set_closure(*args)
    O.set_closure(*args)
    G.set_closure_mode(_lib.CLOSURE_FP32)
    O.set_closure_mode(_lib.CLOSURE_FP32)
    G.step(25)
    O.run(25)
    s = (slice(None),) + (slice(1, -1),) * 3
    assert np.array_equal(G.f()[s], O.f()[s].astype(np.float32))
    assert np.array_equal(G.velocity()[s], O.velocity()[s].astype(np.float32))
    G.set_closure(None)  # plain LBGK again
    O.set_closure(None)
    G.step(5)
    O.run(5)
    assert np.array_equal(G.f()[s], O.f()[s].astype(np.float32))


def test_closure_step_cavity_bitwise():
    from gpu_common import add_bc, compare, pair, run
    G, O = pair((12, 11, 10), 19, nu=0.005)
    G.init_equilibrium(1.0, (0, 0, 0))
    O.init_f(1.0, (0, 0, 0))
    args = closure_args(11)
    G.set_closure(*args)
    O.set_closure(*args)
    G.set_closure_mode(_lib.CLOSURE_FP32)
    O.set_closure_mode(_lib.CLOSURE_FP32)
    bcs = [add_bc(G, O, _lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.08, 0, 0)), add_bc(G, O, _lib.BC_NOSLIP, WALLS)]
    run(G, O, 120, bcs)
    d = compare(G, O)
    assert d["rho"] == 0.0 and d["u"] == 0.0 and d["f"] == 0.0, d


@pytest.mark.parametrize("size", [(10, 9, 8), (33, 17, 12)])
def test_closure_matrix_cores_match_oracle(size):
    """The BF16 matrix-core evaluation (the default) against the oracle's BF16
    model: same roundings, sums in a different order, so a few ulp apart."""
    G, O = periodic_pair(size)
    args = closure_args(5)
    G.set_closure(*args)
    O.set_closure(*args)  # both default to the BF16 evaluation
    s = (slice(None),) + (slice(1, -1),) * 3
    G.step(1)
    O.run(1)
    d1 = np.abs(G.f()[s].astype(np.float64) - O.f()[s]).max()
    G.step(20)
    O.run(20)
    d = np.abs(G.f()[s].astype(np.float64) - O.f()[s]).max()
    print("closure BF16 MFMA vs oracle: max|df| after 1 step %.2e, after 21 steps %.2e" % (d1, d))
    assert d1 < 2e-6 and d < 2e-5
    # the FP32 evaluation is a different rounding of the same network
    G2, _ = periodic_pair(size)
    G2.set_closure(*args)
    G2.set_closure_mode(_lib.CLOSURE_FP32)
    G2.step(21)
    assert 0 < np.abs(G2.f()[s].astype(np.float64) - G.f()[s]).max() < 1e-3


def test_closure_rejects_other_lattices():
    L = asl_amd.Lattice((8, 8, 8), 27, nu=0.1)
    with pytest.raises(asl_amd.ASLError):
        L.set_closure(*closure_args())
    L = asl_amd.Lattice((8, 8, 8), 19, dtype="f64", nu=0.1)
    with pytest.raises(asl_amd.ASLError):
        L.set_closure(*closure_args())


def data(n=3000, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 7)).astype(np.float32)
    t = (np.maximum(x[:, 0] + 0.5 * x[:, 3], 0) + 0.1 * rng.standard_normal(n)).astype(np.float32)
    return x, t


def test_gradient_matches_numpy():
    x, t = data()
    p = mlp.init_params(2)
    m = asl_amd.Mlp(p)
    m.set_data(x, t)
    for first, count in ((0, 3000), (17, 1000), (2999, 1)):
        loss, g = m.gradient(first, count)
        lr, gr = mlp.loss_and_grad(p.astype(np.float64), x[first:first + count], t[first:first + count])
        assert abs(loss - lr) <= 1e-5 * abs(lr)
        assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max()


def batch_indices(seed, step, batch, n):
    """The trainer's counter-based sample choice (splitmix64) for Adam step `step` (from 1)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    G = 0x9E3779B97F4A7C15

    def mix(z):
        z = np.asarray(z, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return (z ^ (z >> np.uint64(31))) & M
    sd = int(mix(np.uint64((seed + G * step) % 2 ** 64)))
    sd = sd or 1
    k = np.arange(batch, dtype=np.uint64) + np.uint64(1)
    with np.errstate(over="ignore"):
        z = np.uint64(sd) + k * np.uint64(G)
    return (mix(z) % np.uint64(n)).astype(np.int64)


def test_adam_steps_match_numpy():
    x, t = data(4000)
    p0 = mlp.init_params(9)
    m = asl_amd.Mlp(p0)
    m.set_data(x, t)
    steps, batch, lr, seed = 4, 512, 3e-3, 12345
    m.train(steps, batch, lr, seed)
    p = p0.astype(np.float64)
    mm, vv = np.zeros_like(p), np.zeros_like(p)
    for s in range(1, steps + 1):
        idx = batch_indices(seed, s, batch, x.shape[0])
        _, g = mlp.loss_and_grad(p, x[idx], t[idx])
        p, mm, vv = mlp.adam_step(p, g, mm, vv, s, lr)
    got = m.params().astype(np.float64)
    # Adam's first steps move each parameter by about lr; a parameter whose
    # gradient is near zero may move either way in FP32 and FP64
    d = np.abs(got - p)
    assert np.mean(d <= 0.02 * lr * steps) > 0.98
    assert d.max() <= 2.0 * lr * steps


def test_training_lowers_loss_on_filtered_fine_data():
    ex = closure.Experiment(n_coarse=16, tau_fine=0.52, seed=3)
    x, t = ex.training_data(coarse_steps_between=6, snapshots=3, log=lambda *a: None)
    assert x.shape == (3 * 16 ** 3, 7) and np.isfinite(x).all() and (t >= -1.5 * ex.nu_c - 1e-9).all()
    params, mu, sc, out, losses = closure.train(x, t, closure.init_params(1), steps=400, batch=2048, lr=3e-3)
    print("closure training loss: first %.4f, last %.4f" % (losses[:10].mean(), losses[-10:].mean()))
    assert losses[-20:].mean() < 0.6 * losses[:5].mean()
    L = ex.coarse()
    L.set_closure(params, mu, sc, out)
    L.step(50)
    assert np.isfinite(L.conservation()).all()
