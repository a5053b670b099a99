"""Visualisation fills of rho / u on the GPU (generateBCNoSlipVel,
generateBCNoSlipRho, generateBCVelocityVel) against the oracle, bit for bit,
on a channel past a sphere after a run of steps.

Compared cells: the guarded boundary nodes of the map (map <= 0 and
> -0.9999) whose fluid neighbours are all computation nodes (their rho / u
are equal in both by the step parity); for the density fill only the nodes
it writes (weight sum above 0.1).  The fills leave the populations alone:
later steps still equal the oracle.
"""
import numpy as np
import pytest

import asl_amd
import oracle
from asl_amd import _lib

pytestmark = pytest.mark.gpu
SIZE = (26, 20, 18)


def sphere_map():
    S = tuple(s + 2 for s in SIZE)
    X, Y, Z = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in S], indexing="ij")
    return (np.sqrt((X - 11.0) ** 2 + (Y - 10.3) ** 2 + (Z - 9.6) ** 2) - 4.2).astype(np.float32)


def channel(q, dtype):
    m = sphere_map()
    G = asl_amd.Lattice(SIZE, q, dtype=dtype, nu=0.04)
    O = oracle.OracleLattice(SIZE, q, fp64=(dtype == "f64"), nu=0.04)
    for L in (G, O):
        L.set_geometry(0.5, (1.0, -2.0, 0.25))
    G.init_equilibrium(1.0, (0.03, 0.0, 0.0))
    O.init_f(1.0, (0.03, 0.0, 0.0))
    spec = [(_lib.BC_CONSTANT_VELOCITY, _lib.X0, 0.0, (0.03, 0.0, 0.0)),
            (_lib.BC_CONSTANT_PRESSURE, _lib.XE, 1.0, None),
            (_lib.BC_NOSLIP, _lib.Y0 | _lib.YE | _lib.Z0 | _lib.ZE, 0.0, None)]
    ids = []
    for kind, faces, p, u in spec:
        ids.append((G.add_slice_bc(kind, faces, p=p, u=u), O.add_slice_bc(kind, faces, p=p, u=u)))
    ids.append((G.add_map_bc(_lib.BC_NOSLIP_MAP, m), O.add_map_bc(O.NOSLIP_MAP, m)))
    for g, o in ids:
        G.execute_bc(g)
        O.execute_bc(o)
    return G, O, m, [o for _, o in ids]


def neighbours(q):
    c, _, _ = oracle.template(q)
    return c


def compare_mask(G, m, q):
    """Guarded nodes whose neighbours with map > 0 are all computation nodes."""
    ct = G.cell_types()
    comp = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    g = (m <= 0) & (m > -0.9999)
    ok = g.copy()
    for ci in neighbours(q)[1:]:
        fl = np.roll(m, tuple(-ci), axis=(0, 1, 2)) > 0
        cn = np.roll(comp, tuple(-ci), axis=(0, 1, 2))
        ok &= ~fl | cn
    ok[[0, -1], :, :] = False
    ok[:, [0, -1], :] = False
    ok[:, :, [0, -1]] = False
    return ok


def density_written(m, q):
    """Nodes whose density-fill weight sum exceeds 0.1 (normal from the map's stencil gradient)."""
    c = neighbours(q)
    gc = {15: (1.0 / 3.0, 1.0 / 24.0), 19: (1.0 / 6.0, 1.0 / 12.0)}[q]
    mm = m.astype(np.float64)
    grad = np.zeros((3,) + m.shape)
    for i, ci in enumerate(c[1:], start=1):
        w = gc[0] if i < 7 else gc[1]
        grad += w * ci.reshape(3, 1, 1, 1) * np.roll(mm, tuple(-ci), axis=(0, 1, 2))
    n = grad / np.sqrt((grad ** 2).sum(0))
    s = np.zeros(m.shape)
    for ci in c[1:]:
        d = ((n * (ci / np.linalg.norm(ci)).reshape(3, 1, 1, 1)).sum(0) + 1.0) ** 2
        s += np.where(np.roll(mm, tuple(-ci), axis=(0, 1, 2)) > 0, d, 0.0)
    return s > 0.1 + 1e-6


@pytest.mark.parametrize("q,dtype", [(19, "f32"), (15, "f32"), (19, "f64")])
def test_fills_match_oracle(q, dtype):
    G, O, m, order = channel(q, dtype)
    G.step(60)
    O.run(60, order)
    mask = compare_mask(G, m, q)
    assert mask.sum() > 100
    vel = G.add_vis(_lib.VIS_NOSLIP_VEL, m)
    rho = G.add_vis(_lib.VIS_NOSLIP_RHO, m)
    G.execute_vis(vel)
    G.execute_vis(rho)
    O.vis_fill(1, m)
    O.vis_fill(2, m)
    cast = np.float64 if dtype == "f64" else np.float32
    ug, uo = G.velocity(), O.velocity().astype(cast)
    rg, ro = G.rho(), O.rho().astype(cast)
    for d in range(3):
        assert np.array_equal(ug[d][mask], uo[d][mask]), d
    dm = mask & density_written(m, q)
    assert dm.sum() > 100
    assert np.array_equal(rg[dm], ro[dm])
    # the mirrored velocity points against the stream next to the sphere
    assert (ug[0][mask] < 0).mean() > 0.5
    # rotating-wall fill at the Block's physical positions
    axis, centre = (0.01, -0.02, 0.015), (6.0, 3.0, 4.0)
    rot = G.add_vis(_lib.VIS_VELOCITY, m, rot=axis + centre)
    G.execute_vis(rot)
    O.vis_fill(3, m, rot=np.array(axis + centre))
    ug, uo = G.velocity(), O.velocity().astype(cast)
    g = (m <= 0) & (m > -0.9999)
    for d in range(3):
        assert np.array_equal(ug[d][g], uo[d][g]), d
    # the populations are untouched: more steps still equal the oracle
    G.step(25)
    O.run(25, order)
    ct = G.cell_types()
    comp = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    assert np.array_equal(G.rho()[comp], O.rho().astype(cast)[comp])


def test_fill_follows_deferred_step_and_rejects_d3q27():
    G, O, m, order = channel(19, "f32")
    G.step(10)
    G.step(1, defer=True)  # what LBGK::execute() does: the macro write waits for a reader
    O.run(11, order)
    vid = G.add_vis(_lib.VIS_NOSLIP_VEL, m)
    G.execute_vis(vid)
    O.vis_fill(1, m)
    mask = compare_mask(G, m, 19)
    assert np.array_equal(G.velocity()[0][mask], O.velocity()[0][mask].astype(np.float32))
    L = asl_amd.Lattice((8, 8, 8), 27, nu=0.1)
    L.init_equilibrium(1.0, (0.0, 0.0, 0.0))
    mm = np.ones((10, 10, 10), dtype=np.float32)
    with pytest.raises(asl_amd.ASLError):
        L.add_vis(_lib.VIS_NOSLIP_RHO, mm)
    L.execute_vis(L.add_vis(_lib.VIS_VELOCITY, mm, u=(0.1, 0.0, 0.0)))
