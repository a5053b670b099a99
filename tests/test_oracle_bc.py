"""Boundary conditions of the oracle: point lists, masks, and the link model.

Pins the oracle's restatement of ASL's slice and map BCs
(src/num/aslBCond.cxx:373-510, src/num/aslLBGKBC.cxx:91-432) against
  * the entry counts of addSlice* (N^2 x (#directions entering) per face),
  * hand-derived masks of a small level-set map,
  * a single tagged population bounced back by a wall,
  * the independent per-link numpy model in tests/linkmodel.py on small
    lattices with every BC kind and overlapping BCs in both execute orders.
"""
import numpy as np
import pytest

from asl_amd import geometry
from linkmodel import (CONSTANT_PRESSURE, CONSTANT_PRESSURE_VELOCITY, CONSTANT_VELOCITY, NOSLIP, NOSLIP_MAP,
                       PRESSURE_VELOCITY_MAP, VELOCITY_MAP, LinkModel)
from oracle import OracleLattice, template

ALL = 63


@pytest.mark.parametrize("q,per_face", [(15, 5), (19, 5), (27, 9)])
def test_slice_entry_counts(q, per_face):
    size = (6, 5, 4)
    L = OracleLattice(size, q, fp64=True)
    faces = [(1, size[1] * size[2]), (2, size[1] * size[2]), (4, size[0] * size[2]), (8, size[0] * size[2]),
             (16, size[0] * size[1]), (32, size[0] * size[1])]
    for bit, area in faces:
        b = L.add_slice_bc(L.NOSLIP, bit)
        assert L.slice_entries(b) == area * per_face
    b = L.add_slice_bc(L.NOSLIP, ALL)
    assert L.slice_entries(b) == per_face * 2 * (size[1] * size[2] + size[0] * size[2] + size[0] * size[1])


def test_map_masks_hand_derived():
    # sphere of radius 1.5 centred on interior cell (3,3,3) of a 6^3 block,
    # normalised: map = clamp((sqrt(d2) - 1.5) / 1.8, -1, 1)
    size = (6, 6, 6)
    m = geometry.normalize(geometry.sphere(size, 1.5, (3.0, 3.0, 3.0)))
    L = OracleLattice(size, 19, fp64=True)
    b = L.add_map_bc(L.NOSLIP_MAP, m)
    masks = L.map_masks(b)
    # ghost-inclusive index I = interior index + 1
    I = np.indices(m.shape)
    d2 = ((I - 4.0) ** 2).sum(0)
    ghost = d2 <= 2.25  # sqrt(d2) <= 1.5
    assert ((masks & 1) != 0).tolist() == ghost.tolist()
    # every ghost node here is within the guard: depth <= 1.5 < 1.8*0.9999
    assert ((masks & 2) != 0).tolist() == ghost.tolist()
    assert ((masks & 4) != 0).tolist() == (~ghost).tolist()
    assert ghost.sum() == 1 + 6 + 12  # centre, faces, edges of the centre cell


def test_guard_excludes_deep_cells():
    # an unnormalised map: depth > 0.9999 fails the guard
    size = (6, 6, 6)
    m = geometry.sphere(size, 2.5, (3.0, 3.0, 3.0))
    L = OracleLattice(size, 19, fp64=True)
    b = L.add_map_bc(L.NOSLIP_MAP, m)
    masks = L.map_masks(b)
    deep = m <= np.float32(-0.9999)
    assert ((masks & 2) != 0).tolist() == ((m <= 0) & ~deep).tolist()
    assert deep.any()


def test_tagged_population_bounces_back():
    # nu = inf: no collision.  A population moving towards the Z0 wall
    # returns reversed one step later (half-way bounce-back).
    size = (4, 4, 4)
    L = OracleLattice(size, 19, fp64=True, nu=float("inf"))
    L.init_f(1.0, (0.0, 0.0, 0.0))
    b = L.add_slice_bc(L.NOSLIP, ALL)
    f = L.f()
    c, _, inv = template(19)
    k = 13  # (0,-1,-1)
    f[k, 2, 2, 1] = 1.25  # interior cell (1,1,0), moving to y-1, z-1
    L.write_f(f)
    L.execute_bc(b)
    L.run(1, [b])
    g = L.f()
    # it streamed into the ghost (2,1,0) and came back into cell (2,2,1) as inv(13)
    assert g[inv[k], 2, 2, 1] == 1.25
    assert np.count_nonzero(g[:, 1:-1, 1:-1, 1:-1] != 1.0) == 1


def _run_both(size, q, nu, steps, setup, order_fn, compressible=True, u0=(0.0, 0.0, 0.0)):
    O = OracleLattice(size, q, fp64=True, nu=nu)
    O.set_compressible(compressible)
    M = LinkModel(size, q, nu, compressible)
    ids_o, ids_m = setup(O), setup(M)
    O.init_f(1.0, u0)
    M.init(1.0, np.asarray(u0))
    order_o, order_m = order_fn(ids_o), order_fn(ids_m)
    for b in order_o:
        O.execute_bc(b)
    O.run(steps, order_o)
    comp = M.run(steps, order_m)
    return O, M, comp


def _compare(O, M, comp, tol=1e-13):
    r = O.rho()
    v = O.velocity()
    assert np.abs(r[comp] - M.rho[comp]).max() < tol
    with np.errstate(invalid="ignore", divide="ignore"):
        vm = M.m / M.rho if M.compressible else M.m
    for d in range(3):
        assert np.abs(v[d][comp] - vm[d][comp]).max() < tol
    f = O.f()
    for i in range(f.shape[0]):
        assert np.abs(f[i][comp] - M.f[i][comp]).max() < tol


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("lid_last", [False, True])
def test_cavity_matches_link_model(q, lid_last):
    size = (5, 4, 6)

    def setup(L):
        lid = L.add_slice_bc(CONSTANT_VELOCITY, 32, u=(0.05, 0.01, 0.0)) if isinstance(L, OracleLattice) else \
            L.add(CONSTANT_VELOCITY, 32, u=(0.05, 0.01, 0.0))
        walls = L.add_slice_bc(NOSLIP, 31) if isinstance(L, OracleLattice) else L.add(NOSLIP, 31)
        return lid, walls

    order = (lambda ids: [ids[1], ids[0]]) if lid_last else (lambda ids: [ids[0], ids[1]])
    O, M, comp = _run_both(size, q, 0.05, 12, setup, order)
    _compare(O, M, comp)


@pytest.mark.parametrize("compressible", [True, False])
def test_channel_sphere_pressure_matches_link_model(compressible):
    size = (10, 7, 7)
    mapv = geometry.normalize(geometry.sphere(size, 1.7, (4.0, 3.0, 3.0)))
    pmap = geometry.normalize(geometry.sphere(size, 1.2, (8.0, 3.0, 3.0)))

    def setup(L):
        o = isinstance(L, OracleLattice)
        return [
            L.add_slice_bc(CONSTANT_PRESSURE_VELOCITY, 1, p=1.02, u=(0.03, 0, 0)) if o else
            L.add(CONSTANT_PRESSURE_VELOCITY, 1, p=1.02, u=(0.03, 0, 0)),
            L.add_slice_bc(CONSTANT_PRESSURE, 2, p=0.99) if o else L.add(CONSTANT_PRESSURE, 2, p=0.99),
            L.add_slice_bc(NOSLIP, 4 | 8 | 16 | 32) if o else L.add(NOSLIP, 4 | 8 | 16 | 32),
            L.add_map_bc(NOSLIP_MAP, mapv) if o else L.add(NOSLIP_MAP, mapv=mapv),
            L.add_map_bc(PRESSURE_VELOCITY_MAP, pmap, p=1.01, u=(0.01, 0.0, 0.02)) if o else
            L.add(PRESSURE_VELOCITY_MAP, mapv=pmap, p=1.01, u=(0.01, 0.0, 0.02)),
        ]

    O, M, comp = _run_both(size, 19, 0.04, 10, setup, lambda ids: list(ids), compressible=compressible,
                           u0=(0.02, 0.0, 0.0))
    assert (~comp[1:-1, 1:-1, 1:-1]).sum() > 10  # the obstacles remove cells
    _compare(O, M, comp)


@pytest.mark.parametrize("rotating", [False, True])
def test_velocity_map_matches_link_model(rotating):
    # BCVelocityMap (src/num/aslLBGKBC.cxx:290-348): rho at the ghost node
    # from its computation-node neighbours, f_i(g) = rho_g + 3 rho_g u.c_i
    size = (9, 8, 8)
    body = geometry.normalize(geometry.sphere(size, 2.2, (4.0, 3.5, 3.5)))
    rot = (0.0, 0.01, 0.02, 4.0, 3.5, 3.5) if rotating else None

    def setup(L):
        o = isinstance(L, OracleLattice)
        return [L.add_slice_bc(NOSLIP, ALL) if o else L.add(NOSLIP, ALL),
                L.add_map_bc(VELOCITY_MAP, body, u=(0.02, -0.01, 0.01), rot=rot) if o else
                L.add(VELOCITY_MAP, mapv=body, u=(0.02, -0.01, 0.01), rot=rot)]

    O, M, comp = _run_both(size, 19, 0.05, 10, setup, lambda ids: list(ids))
    _compare(O, M, comp)
    assert np.abs(O.velocity()[:, comp]).max() > 1e-3


def _casing(size, radius):
    """Fluid inside a z-cylinder, map = (radius - r) / 3 clipped to [-1, 1]: a
    guard band three cells thick, so no cell beyond it neighbours the fluid."""
    S = tuple(n + 2 for n in size)
    x = np.arange(S[0], dtype=np.float64)[:, None, None] - (S[0] - 1) / 2.0
    y = np.arange(S[1], dtype=np.float64)[None, :, None] - (S[1] - 1) / 2.0
    m = (radius - np.sqrt(x * x + y * y)) / 3.0 + np.zeros((1, 1, S[2]))
    return np.clip(m, -1.0, 1.0).astype(np.float32)


def test_velocity_map_skip_outer_option():
    # Extension (opt-in defined value): velocity-map density averages leave
    # out neighbours in the outer ghost layer.  On a casing whose boundary
    # nodes in the first and last z layer have such neighbours (ASL reads the
    # values the Z0 / ZE pressure BC left there) the option changes the run;
    # where no outer neighbour is involved it changes nothing, bit for bit.
    size = (14, 14, 6)
    m = _casing(size, 4.5)
    inner = (slice(1, -1),) * 3
    fluid = m[inner] > 0
    res = []
    for skip in (False, True):
        O = OracleLattice(size, 15, fp64=False, nu=0.05)
        O.set_vmap_skip_outer(skip)
        O.init_f(1.0, (0.0, 0.0, 0.0))
        ids = [O.add_map_bc(VELOCITY_MAP, m, u=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.01, 7.5, 7.5, 3.5)),
               O.add_slice_bc(CONSTANT_PRESSURE, 16 | 32, p=1.0)]
        for b in ids:
            O.execute_bc(b)
        O.run(6, ids)
        r = O.rho()[inner][fluid]
        assert np.isfinite(r).all()
        res.append(r)
    assert not np.array_equal(res[0], res[1])
    size = (9, 8, 8)
    body = geometry.normalize(geometry.sphere(size, 2.2, (4.0, 3.5, 3.5)))
    out = []
    for skip in (False, True):
        O = OracleLattice(size, 19, fp64=False, nu=0.05)
        O.set_vmap_skip_outer(skip)
        O.init_f(1.0, (0.0, 0.0, 0.0))
        ids = [O.add_slice_bc(NOSLIP, ALL), O.add_map_bc(VELOCITY_MAP, body, u=(0.02, -0.01, 0.01))]
        for b in ids:
            O.execute_bc(b)
        O.run(10, ids)
        out.append(O.f())
    assert np.array_equal(out[0], out[1], equal_nan=True)
