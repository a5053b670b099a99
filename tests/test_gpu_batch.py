"""Batches: many independent lattices of uneven sizes in one launch per step.

Each lattice of a batch must end exactly where it ends when stepped alone
(same kernel body, same operation order): bit for bit, with and without the
hipGraph-captured chunks, including graph replays separated by eager work;
one member is also checked against the CPU oracle.
"""
import numpy as np
import pytest

import asl_amd
import oracle
from asl_amd import _lib

pytestmark = pytest.mark.gpu
WALLS = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0


def cavity(n, tau, lid=0.05):
    L = asl_amd.Lattice((n, n + 1, n + 2), 19, nu=(tau - 0.5) / 3)
    L.init_equilibrium(1.0, (0, 0, 0))
    a = L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(lid, 0, 0))
    b = L.add_slice_bc(_lib.BC_NOSLIP, WALLS)
    L.execute_bc(a)
    L.execute_bc(b)
    return L


SPECS = [(9, 0.55), (14, 0.8), (21, 0.62), (6, 0.95), (17, 0.51)]


def interior(a):
    return a[(slice(None),) * (a.ndim - 3) + (slice(1, -1),) * 3]


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_lattices(graph):
    steps = 237
    solo = [cavity(n, t) for n, t in SPECS]
    for L in solo:
        L.step(steps)
    members = [cavity(n, t) for n, t in SPECS]
    B = asl_amd.Batch(members)
    B.step(steps, graph=graph)
    for L, R in zip(members, solo):
        assert np.array_equal(interior(L.f()), interior(R.f()))
        assert np.array_equal(interior(L.rho()), interior(R.rho()))
        assert np.array_equal(interior(L.velocity()), interior(R.velocity()))


def test_graph_replays_interleaved_with_eager_work():
    members = [cavity(n, t) for n, t in SPECS]
    solo = [cavity(n, t) for n, t in SPECS]
    B = asl_amd.Batch(members)
    total = 0
    for chunk in (201, 7, 300, 101):  # graphs of 100 steps replayed with eager steps between
        B.step(chunk, graph=True)
        total += chunk
        for L in members:  # eager host work between replays
            L.conservation()
    for L in solo:
        L.step(total)
    for L, R in zip(members, solo):
        assert np.array_equal(interior(L.f()), interior(R.f()))


def test_batch_member_matches_oracle():
    n, tau, steps = 12, 0.7, 150
    members = [cavity(m, t) for m, t in SPECS[:2]] + [cavity(n, tau)]
    B = asl_amd.Batch(members)
    B.step(steps, graph=True)
    O = oracle.OracleLattice((n, n + 1, n + 2), 19, nu=(tau - 0.5) / 3)
    O.init_f(1.0, (0, 0, 0))
    ids = [O.add_slice_bc(O.CONSTANT_VELOCITY, 32, u=(0.05, 0, 0)), O.add_slice_bc(O.NOSLIP, 31)]
    for b in ids:
        O.execute_bc(b)
    O.run(steps, ids)
    G = members[-1]
    assert np.array_equal(interior(G.rho()), interior(O.rho()).astype(np.float32))
    assert np.array_equal(interior(G.velocity()), interior(O.velocity()).astype(np.float32))


def test_batch_workgroup_size_keeps_results():
    solo = [cavity(n, t) for n, t in SPECS]
    for L in solo:
        L.step(130)
    members = [cavity(n, t) for n, t in SPECS]
    B = asl_amd.Batch(members)
    B.set_tuning(_lib.TUNE_LDS, 0)  # the direct kernel (FP32 batches default to LDS tiles)
    B.set_tuning(_lib.TUNE_BLOCK, 256)
    B.step(30, graph=True)
    B.set_tuning(_lib.TUNE_LDS, 1)
    B.step(30, graph=True)
    assert B.autotune(5) in (256, 512, 1024)  # (tiles + 3 workgroup sizes) x 6 real steps
    B.step(130 - 60 - 24, graph=True)
    for L, R in zip(members, solo):
        assert np.array_equal(interior(L.f()), interior(R.f()))


@pytest.mark.parametrize("group_mb", [1, 3])
def test_lattice_groups_keep_results(group_mb):
    """Graph chunks that step groups of lattices one group at a time (here
    groups of one to a few of these small lattices) end bit for bit where the
    single lattices end."""
    steps = 300
    solo = [cavity(n, t) for n, t in SPECS]
    for L in solo:
        L.step(steps)
    members = [cavity(n, t) for n, t in SPECS]
    B = asl_amd.Batch(members)
    B.set_tuning(_lib.TUNE_GROUP_MB, group_mb)
    B.step(steps, graph=True)
    for L, R in zip(members, solo):
        assert np.array_equal(interior(L.f()), interior(R.f()))
        assert np.array_equal(interior(L.rho()), interior(R.rho()))
    assert B.autotune_groups(1) in (0, 64, 128, 160, 192, 224, 256, 384)
