"""The oracle's ComputeSurfaceForceMap / ComputeSurfaceFluxMap restatement.

Pinned by an independent numpy evaluation written on 3-D coordinates (so a
wrong shift sign, direction or inverse fails), by the rest state (no force,
no flux) and by the direction of the drag on a body in a stream.
"""
import numpy as np

from asl_amd import geometry
from oracle import OracleLattice, template


def numpy_surface(f, mapv, c, w, inv):
    """Force (3,)+S and flux S from f (q,)+S on interior cells, neighbours by 3-D shift."""
    q = f.shape[0]
    S = mapv.shape
    force = np.zeros((3,) + S)
    flux = np.zeros(S)
    X = np.indices(S)
    for i in range(1, q):
        Y = [X[d] + c[i][d] for d in range(3)]
        inside = (Y[0] >= 0) & (Y[0] < S[0]) & (Y[1] >= 0) & (Y[1] < S[1]) & (Y[2] >= 0) & (Y[2] < S[2])
        Yc = [np.clip(Y[d], 0, S[d] - 1) for d in range(3)]
        my = np.where(inside, mapv[Yc[0], Yc[1], Yc[2]], -0.9)
        fI = f[inv[i]][Yc[0], Yc[1], Yc[2]]
        f0 = f[i]
        b0 = (mapv <= 0) & (my > 0)
        bi = (mapv > 0) & (my <= 0)
        for d in range(3):
            force[d] += np.where(b0, (f0 + fI - 2) * w[i] * c[i][d], 0.0)
            force[d] += np.where(bi, (-f0 - fI + 2) * w[i] * c[i][d], 0.0)
        flux += np.where(b0, (f0 - fI) * w[i], 0.0) + np.where(bi, (fI - f0) * w[i], 0.0)
    keep = np.abs(mapv) < 0.9999
    return force * keep, flux * keep


def channel(size=(14, 10, 9), steps=12):
    mapv = geometry.normalize(geometry.sphere(size, 2.5, (6.0, 5.0, 4.5)))
    O = OracleLattice(size, 19, fp64=True, nu=0.05)
    O.init_f(1.0, (0.03, 0.0, 0.0))
    bcs = [O.add_slice_bc(O.CONSTANT_VELOCITY, 1, u=(0.03, 0, 0)), O.add_slice_bc(O.CONSTANT_PRESSURE, 2, p=1.0),
           O.add_slice_bc(O.NOSLIP, 4 | 8 | 16 | 32), O.add_map_bc(O.NOSLIP_MAP, mapv)]
    for b in bcs:
        O.execute_bc(b)
    O.run(steps, bcs)
    return O, mapv


def test_surface_matches_independent_numpy():
    O, mapv = channel()
    c, w, inv = template(19)
    f = O.f()
    ref_force, ref_flux = numpy_surface(f, mapv, c, w, inv)
    inner = (slice(1, -1),) * 3  # away from the block edge (ASL wraps the linear index there)
    got_force, got_flux = O.surface_force(mapv), O.surface_flux(mapv)
    assert np.abs(got_force[(slice(None),) + inner] - ref_force[(slice(None),) + inner]).max() < 1e-15
    assert np.abs(got_flux[inner] - ref_flux[inner]).max() < 1e-15
    assert np.abs(got_force).max() > 1e-4


def test_rest_state_has_no_force_or_flux():
    size = (8, 8, 8)
    mapv = geometry.normalize(geometry.sphere(size, 2.0, (4.5, 4.5, 4.5)))
    O = OracleLattice(size, 19, fp64=True, nu=0.1)
    O.init_f(1.0, (0.0, 0.0, 0.0))
    b = O.add_map_bc(O.NOSLIP_MAP, mapv)
    w = O.add_slice_bc(O.NOSLIP, 63)
    O.execute_bc(w)
    O.execute_bc(b)
    O.run(3, [w, b])
    inner = (slice(1, -1),) * 3  # the outer layer reads beyond the block (undefined in ASL too)
    assert np.abs(O.surface_force(mapv)[(slice(None),) + inner]).max() < 1e-14
    assert np.abs(O.surface_flux(mapv)[inner]).max() < 1e-14
    assert ((np.abs(mapv) < 0.9999) & (mapv <= 0))[inner].any()  # there is a surface


def test_drag_points_downstream():
    """ASL's field is the momentum the links give the fluid: with a bounce-back
    link, (f_i(g) + f_inv(y) - 2) w c_i = 2 (f_inv(y) - 1) w c_i, minus what the
    population arriving at the body hands it; so it points against the stream."""
    O, mapv = channel(steps=40)
    F = O.surface_force(mapv)
    inner = (slice(1, -1),) * 3
    ghost = (mapv <= 0)[inner]
    fx, fy = F[0][inner][ghost].sum(), F[1][inner][ghost].sum()
    assert fx < 0
    assert abs(fy) < 0.2 * abs(fx)
