"""Self-generated regression fixtures (tests/golden/lbgk_fixtures.npz, written
by tests/golden/make_fixtures.py from the CPU oracle; ASL ships no LBGK test
vectors).  The oracle must keep reproducing them bit for bit, and so must the
GPU (FP64 cavities, FP32 sphere channel), on interior / computation nodes."""
import os

import numpy as np
import pytest

import oracle
from asl_amd import _lib

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbgk_fixtures.npz")
INNER = (slice(1, -1),) * 3


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX) as z:  # allow_pickle stays False
        return {k: z[k] for k in z.files}


def oracle_cavity(q, steps):
    O = oracle.OracleLattice((8, 8, 8), q, fp64=True, nu=1.0 / 30.0)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 32, u=(0.05, 0.0, 0.0)), O.add_slice_bc(O.NOSLIP, 31)]
    for b in ids:
        O.execute_bc(b)
    O.run(steps, ids)
    return O


@pytest.mark.parametrize("q", [15, 19, 27])
def test_oracle_reproduces_cavity_fixtures(fx, q):
    O = oracle_cavity(q, 10)
    assert np.array_equal(O.f()[(slice(None),) + INNER], fx["cavity8_q%d_f10" % q][(slice(None),) + INNER])
    O = oracle_cavity(q, 100)
    assert np.array_equal(O.rho()[INNER], fx["cavity8_q%d_rho100" % q][INNER])
    assert np.array_equal(O.velocity()[(slice(None),) + INNER], fx["cavity8_q%d_u100" % q][(slice(None),) + INNER])
    # a moving lid: the fixture holds a flow, not a rest state
    assert np.abs(fx["cavity8_q%d_u100" % q][(slice(None),) + INNER]).max() > 1e-3


def test_oracle_reproduces_channel_fixture(fx):
    body = fx["channel_q19_f32_map"]
    O = oracle.OracleLattice((16, 8, 8), 19, fp64=False, nu=0.02)
    O.init_f(1.0, (0.04, 0.0, 0.0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 1, u=(0.04, 0.0, 0.0)), O.add_slice_bc(O.CONSTANT_PRESSURE, 2, p=1.0),
           O.add_slice_bc(O.NOSLIP, 4 | 8 | 16 | 32), O.add_map_bc(O.NOSLIP_MAP, body)]
    for b in ids:
        O.execute_bc(b)
    O.run(50, ids)
    fluid = body[INNER] > 0
    assert np.array_equal(O.rho()[INNER][fluid].astype(np.float32), fx["channel_q19_f32_rho50"][INNER][fluid])


@pytest.mark.gpu
@pytest.mark.parametrize("q", [15, 19, 27])
def test_gpu_reproduces_cavity_fixtures(fx, q):
    import asl_amd
    for steps in (10, 100):
        G = asl_amd.Lattice((8, 8, 8), q, dtype="f64", nu=1.0 / 30.0)
        G.init_equilibrium(1.0, (0.0, 0.0, 0.0))
        for b in (G.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)),
                  G.add_slice_bc(_lib.BC_NOSLIP, 31)):
            G.execute_bc(b)
        G.step(steps)
        if steps == 10:
            assert np.array_equal(G.f()[(slice(None),) + INNER], fx["cavity8_q%d_f10" % q][(slice(None),) + INNER])
        else:
            assert np.array_equal(G.rho()[INNER], fx["cavity8_q%d_rho100" % q][INNER])
            assert np.array_equal(G.velocity()[(slice(None),) + INNER],
                                  fx["cavity8_q%d_u100" % q][(slice(None),) + INNER])


@pytest.mark.gpu
def test_gpu_reproduces_channel_fixture(fx):
    import asl_amd
    body = fx["channel_q19_f32_map"]
    G = asl_amd.Lattice((16, 8, 8), 19, dtype="f32", nu=0.02)
    G.init_equilibrium(1.0, (0.04, 0, 0))
    for b in (G.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.04, 0, 0)),
              G.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
              G.add_slice_bc(_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
              G.add_map_bc(_lib.BC_NOSLIP_MAP, body)):
        G.execute_bc(b)
    G.step(50)
    ct = G.cell_types()
    comp = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    assert comp.sum() > 500
    assert np.array_equal(G.rho()[comp], fx["channel_q19_f32_rho50"][comp])
    assert np.array_equal(G.velocity()[:, comp], fx["channel_q19_f32_u50"][:, comp])
