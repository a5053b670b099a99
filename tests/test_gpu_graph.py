"""Single lattices stepped through hipGraph-captured chunks of 100 steps.

A replayed chunk launches the same kernels with the same arguments as the
eager path, so the result must be bit for bit the eager one: checked after
every call of a sequence that mixes replays, eager steps, deferred steps,
host reads and BC-order changes (which rebuild the tables the graphs use).
The velocity-map link fill is part of the captured chunk and is covered by a
rotating-cylinder lattice.
"""
import numpy as np
import pytest

import asl_amd
from asl_amd import _lib, geometry

pytestmark = pytest.mark.gpu
WALLS = _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0


def cavity(n=20):
    L = asl_amd.Lattice((n, n - 3, n + 2), 19, nu=(0.6 - 0.5) / 3)
    L.init_equilibrium(1.0, (0, 0, 0))
    bcs = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0, 0)),
           L.add_slice_bc(_lib.BC_NOSLIP, WALLS)]
    for b in bcs:
        L.execute_bc(b)
    return L, bcs


def rotor():
    size = (20, 16, 14)
    body = geometry.normalize(geometry.sphere(size, 3.5, (9.0, 7.5, 6.5)))
    L = asl_amd.Lattice(size, 19, nu=0.05)
    L.init_equilibrium(1.0, (0, 0, 0))
    bcs = [L.add_slice_bc(_lib.BC_NOSLIP, 63),
           L.add_map_bc(_lib.BC_VELOCITY_MAP, body, u=(0.02, 0, 0.01), rot=(0, 0, 0.01, 9.0, 7.5, 6.5))]
    for b in bcs:
        L.execute_bc(b)
    return L, bcs


def same(A, B):
    s = (slice(None),) + (slice(1, -1),) * 3
    assert np.array_equal(A.f()[s], B.f()[s])
    assert np.array_equal(A.rho()[s[1:]], B.rho()[s[1:]])
    assert np.array_equal(A.velocity()[s], B.velocity()[s])


@pytest.mark.parametrize("make", [cavity, rotor])
def test_graph_chunks_equal_eager_steps(make):
    G, _ = make()
    E, _ = make()
    for n in (1, 250, 100, 37, 201):  # first call includes the raw first step
        G.step(n, graph=True)
        E.step(n)
        same(G, E)


def test_graph_with_deferred_steps_and_order_change():
    G, gb = cavity()
    E, eb = cavity()
    for n in (130, 5, 220):
        G.step(n, defer=True, graph=True)
        E.step(n, defer=True)
        same(G, E)
    # swap the execute order: no-slip now loses the lid edges; tables and graphs are rebuilt
    for L, b in ((G, gb), (E, eb)):
        L.execute_bc(b[1])
        L.execute_bc(b[0])
    G.step(300, graph=True)
    E.step(300)
    same(G, E)


def test_timed_graph_steps():
    G, _ = cavity()
    E, _ = cavity()
    ms_total, ms_kernel = G.time_steps(400, graph=True)
    E.step(400, write_macro=False)
    assert ms_total > 0 and ms_kernel == 0
    G.step(1)
    E.step(1)
    same(G, E)


def test_launch_tuning_and_autotune_keep_results():
    """Workgroup size and chunk order change the launch only, never the result."""
    ref, _ = cavity()
    ref.step(120)
    for lds, block, remap in ((0, 256, 0), (0, 512, 16), (0, 64, 3), (0, 1024, 4096), (1, 512, 0), (1, 512, 4096),
                              (2, 1024, 0), (2, 64, 16)):
        L, _ = cavity()
        L.set_tuning(_lib.TUNE_LDS, lds)
        L.set_tuning(_lib.TUNE_BLOCK, block)
        L.set_tuning(_lib.TUNE_XCD_REMAP, remap)
        L.step(70, graph=True)
        L.step(50)
        same(L, ref)
    L, _ = cavity()
    assert L.tuning() == (512, 4)  # the defaults for a small lattice: direct, workgroups of 512
    # candidates: tiles on every step x 6 XCD runs; tiles on even steps x 6 XCD
    # runs x 3 odd-step workgroups; direct with cells per lane 1 and 2 (nz = 22
    # is not a multiple of 4) x 6 XCD runs x 3 workgroup sizes; 10 real steps each
    ncand = 6 + 6 * 3 + 2 * 6 * 3
    ref.step(10 * ncand)  # ref: 120 + 10 ncand steps; L: 10 ncand + 120
    best = L.autotune(10)
    assert best[0] in (256, 512, 1024) and best[1] in (0, 2, 4, 8, 16, 64)
    assert L.tuning_value(_lib.TUNE_CPL) in (1, 2)
    assert L.tuning() == best
    L.step(120)
    same(L, ref)
    with pytest.raises(asl_amd.ASLError):
        L.set_tuning(_lib.TUNE_BLOCK, 100)


def test_graph_follows_flag_changes():
    """A captured chunk names the kernel variant and argument slot of the
    flags it was captured with; switching the rotating frame, the
    compressibility or the turbulence limiter between replays must re-capture.
    Checked against eager steps of a twin lattice."""
    A, _ = cavity()
    B, _ = cavity()
    for setter, value in ((None, None), ("set_omega", (0.0, 0.0, 0.01)), ("set_compressible", False),
                          ("set_turbulence", True), ("set_omega", None)):
        if setter:
            getattr(A, setter)(value)
            getattr(B, setter)(value)
        A.step(200, graph=True)
        B.step(200)
        same(A, B)


@pytest.mark.parametrize("periodic", [False, True])
def test_wide_addressing_variant_matches(periodic):
    """The 64-bit addressing variant (used by lattices over 4 GiB of
    populations), forced on a small lattice, ends bit for bit where the 32-bit
    one does, in eager steps, graph chunks and the closure's feature kernel."""
    size = (18, 14, 12)
    body = geometry.normalize(geometry.sphere(size, 3.0, (8.0, 6.5, 5.5)))
    runs = []
    for wide in (0, 1):
        L = asl_amd.Lattice(size, 19, nu=0.04, periodic=periodic)
        L.set_tuning(_lib.TUNE_WIDE, wide)
        rng = np.random.default_rng(2)
        u = (0.03 * rng.standard_normal((3,) + L.S)).astype(np.float32)
        L.init_equilibrium(1.0, u)
        if not periodic:
            for b in (L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0, 0)),
                      L.add_slice_bc(_lib.BC_NOSLIP, _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
                      L.add_map_bc(_lib.BC_NOSLIP_MAP, body)):
                L.execute_bc(b)
        L.step(57)
        L.step(200, graph=True)
        runs.append((L.f(), L.rho(), L.velocity(), L.closure_features() if periodic else None))
    s = (slice(None),) + (slice(1, -1),) * 3
    assert np.array_equal(runs[0][0][s], runs[1][0][s])
    assert np.array_equal(runs[0][1][s[1:]], runs[1][1][s[1:]])
    assert np.array_equal(runs[0][2][s], runs[1][2][s])
    if periodic:
        assert np.array_equal(runs[0][3], runs[1][3])
