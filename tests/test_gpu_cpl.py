"""Runs of 2 or 4 cells per lane (the step kernel's vector variant) and
LDS-staged tiles of 256 cells end bit for bit where one cell per lane ends:
same addresses, same link resolution, same collision per cell.  Covered:
every slice BC kind on every face in both execute orders, map BCs (no-slip,
pressure, velocity with its boundary-link kernel), the rotating frame,
incompressible mode, D3Q15/19/27, all storage types, graph replays, sizes
whose nz does not divide by the run, and planes of several tiles whose
boundaries cut rows (bounce-back partners outside the tile).
"""
import numpy as np
import pytest

import asl_amd
from asl_amd import _lib

pytestmark = pytest.mark.gpu
INNER = (slice(1, -1),) * 3


def state(L):
    return L.f()[(slice(None),) + INNER], L.rho()[INNER], L.velocity()[(slice(None),) + INNER]


def same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


def sphere(S, c, r):
    X, Y, Z = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in S], indexing="ij")
    return (np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r).astype(np.float32)


def build(size, q, dtype, cpl, case):
    """cpl: cells per lane (1: the direct kernel), "lds" for LDS-staged tiles
    on every step, "lds_even" for tiles on even steps only."""
    L = asl_amd.Lattice(size, q, dtype=dtype, nu=0.05)
    if cpl == "lds":
        L.set_tuning(_lib.TUNE_LDS, 1)
    elif cpl == "lds_even":
        L.set_tuning(_lib.TUNE_LDS, 2)
    else:
        L.set_tuning(_lib.TUNE_LDS, 0)
        L.set_tuning(_lib.TUNE_CPL, cpl)
    S = L.S
    if case == "slices":
        L.init_equilibrium(1.0, (0.02, -0.01, 0.005))
        ids = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE | _lib.X0, u=(0.05, 0.01, 0.0)),
               L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.01),
               L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE_VELOCITY, _lib.Y0, p=0.99, u=(0.0, 0.02, 0.01)),
               L.add_slice_bc(_lib.BC_NOSLIP, _lib.YE | _lib.Z0 | _lib.X0)]
        for i in ids:
            L.execute_bc(i)
    elif case == "slices_reversed":
        L.init_equilibrium(1.0, (0.0, 0.0, 0.0))
        ids = [L.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
               L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.ZE, u=(0.05, 0.0, 0.0))]
        for i in ids:
            L.execute_bc(i)
    elif case == "maps":
        L.set_geometry(0.5, (1.0, 2.0, -1.0))
        L.init_equilibrium(1.0, (0.03, 0.0, 0.0))
        m1 = sphere(S, (S[0] * 0.3, S[1] * 0.5, S[2] * 0.45), 2.2)
        m2 = sphere(S, (S[0] * 0.7, S[1] * 0.45, S[2] * 0.55), 1.6)
        ids = [L.add_slice_bc(_lib.BC_CONSTANT_VELOCITY, _lib.X0, u=(0.03, 0.0, 0.0)),
               L.add_slice_bc(_lib.BC_CONSTANT_PRESSURE, _lib.XE, p=1.0),
               L.add_slice_bc(_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
               L.add_map_bc(_lib.BC_NOSLIP_MAP, m1),
               L.add_map_bc(_lib.BC_VELOCITY_MAP, m2, rot=(0.0, 0.0, 0.01, 3.0, 4.0, 2.5))]
        for i in ids:
            L.execute_bc(i)
    elif case == "frame":
        L.set_omega((0.001, -0.002, 0.003))
        L.set_compressible(False)
        L.init_equilibrium(1.0, (0.01, 0.02, 0.0))
        m = sphere(S, (S[0] * 0.5, S[1] * 0.5, S[2] * 0.5), 2.2)
        ids = [L.add_slice_bc(_lib.BC_NOSLIP, _lib.X0 | _lib.XE | _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE),
               L.add_map_bc(_lib.BC_PRESSURE_VELOCITY_MAP, m, p=1.02, u=(0.0, 0.01, 0.0))]
        for i in ids:
            L.execute_bc(i)
    return L


@pytest.mark.parametrize("case", ["slices", "slices_reversed", "maps", "frame"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_runs_equal_single_cells(case, q):
    size = (20, 11, 12) if case == "maps" else (11, 9, 12)
    ref = build(size, q, "f32", 1, case)
    ref.step(41)
    want = state(ref)
    for cpl in (2, 4, "lds", "lds_even"):
        L = build(size, q, "f32", cpl, case)
        if cpl == "lds":
            assert L.tuning_value(_lib.TUNE_LDS) == 1 and L.tuning()[0] == 256
        elif cpl == "lds_even":
            assert L.tuning_value(_lib.TUNE_LDS) == 2 and L.tuning()[0] == 512
        else:
            assert L.tuning_value(_lib.TUNE_CPL) == cpl
        L.step(41)
        assert same(state(L), want), (case, q, cpl)


@pytest.mark.parametrize("case", ["slices", "maps", "frame"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_tiles_across_rows_equal_single_cells(case, q):
    # planes of 23 rows of 32 cells: three tiles per plane (256, 256, 224 cells)
    size = (20, 23, 30)
    ref = build(size, q, "f32", 1, case)
    ref.step(23)
    for v in ("lds", "lds_even"):
        L = build(size, q, "f32", v, case)
        L.step(23)
        assert same(state(L), state(ref)), (case, q, v)


@pytest.mark.parametrize("dtype", ["f64", "bf16", "f16"])
def test_runs_equal_single_cells_storage(dtype):
    size = (20, 11, 8)
    ref = build(size, 19, dtype, 1, "maps")
    ref.step(33)
    want = state(ref)
    for cpl in (2, 4, "lds"):
        L = build(size, 19, dtype, cpl, "maps")
        L.step(33)
        assert same(state(L), want), (dtype, cpl)


def test_runs_under_graph_replay_and_tuning_changes():
    size = (20, 11, 16)
    ref = build(size, 19, "f32", 1, "maps")
    ref.step(257)
    L = build(size, 19, "f32", 4, "maps")
    L.step(120, graph=True)
    L.set_tuning(_lib.TUNE_CPL, 2)
    L.step(37, graph=True)
    L.set_tuning(_lib.TUNE_BLOCK, 256)
    L.set_tuning(_lib.TUNE_CPL, 0)
    L.step(100, graph=True)
    assert same(state(L), state(ref))
    L.set_tuning(_lib.TUNE_LDS, 1)
    assert L.tuning_value(_lib.TUNE_LDS) == 1
    L.step(131, graph=True)
    L.set_tuning(_lib.TUNE_LDS, 2)
    L.step(77, graph=True)
    L.set_tuning(_lib.TUNE_LDS, 0)
    L.step(20, graph=True)
    ref.step(228)
    assert same(state(L), state(ref))


def test_run_length_follows_nz():
    for nz, want in ((12, 4), (10, 2), (9, 1)):
        L = asl_amd.Lattice((6, 5, nz), 19, nu=0.1)
        # the default for a small FP32 lattice: direct, one cell per lane
        assert L.tuning_value(_lib.TUNE_CPL) == 1 and L.tuning_value(_lib.TUNE_LDS) == 0
        L.set_tuning(_lib.TUNE_CPL, 4)
        assert L.tuning_value(_lib.TUNE_CPL) == want
        L.set_tuning(_lib.TUNE_CPL, 2)
        assert L.tuning_value(_lib.TUNE_CPL) == (2 if nz % 2 == 0 else 1)
    L = asl_amd.Lattice((6, 5, 8), 19, nu=0.1, periodic=True)
    L.set_tuning(_lib.TUNE_CPL, 4)
    assert L.tuning_value(_lib.TUNE_CPL) == 1  # periodic wrap: one cell per lane
    with pytest.raises(asl_amd.ASLError):
        L.set_tuning(_lib.TUNE_CPL, 3)
    L.set_tuning(_lib.TUNE_LDS, 1)
    assert L.tuning_value(_lib.TUNE_LDS) == 0  # periodic wrap: direct
    with pytest.raises(asl_amd.ASLError):
        L.set_tuning(_lib.TUNE_LDS, -1)
    L = asl_amd.Lattice((6, 5, 12), 19, nu=0.1)
    L.set_tuning(_lib.TUNE_CPL, 4)  # runs take precedence over tiles
    assert L.tuning_value(_lib.TUNE_LDS) == 0 and L.tuning_value(_lib.TUNE_CPL) == 4
    L.set_tuning(_lib.TUNE_CPL, 0)
    assert L.tuning_value(_lib.TUNE_LDS) == 0 and L.tuning_value(_lib.TUNE_CPL) == 1
    L.set_tuning(_lib.TUNE_LDS, 1)
    assert L.tuning_value(_lib.TUNE_LDS) == 1 and L.tuning() == (256, 4)
    L.set_tuning(_lib.TUNE_LDS, 2)
    assert L.tuning() == (512, 4)
    L.set_tuning(_lib.TUNE_LDS, 0)
    assert L.tuning() == (512, 4)
    with pytest.raises(asl_amd.ASLError):
        L.set_tuning(_lib.TUNE_LDS, 3)
    # automatic tiles: FP32 lattices over 256 MiB of populations (19 x 4 B x 3.7 M cells)
    for dtype, want in (("f32", 1), ("f16", 0), ("bf16", 0), ("f64", 0)):
        L = asl_amd.Lattice((160, 150, 150), 19, dtype=dtype, nu=0.1)
        assert L.tuning_value(_lib.TUNE_LDS) == want, dtype
        assert L.tuning() == ((256, 4) if want else (512, 4 if dtype in ("f32", "f64") else 64))
        del L
