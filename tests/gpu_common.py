"""Shared setup for the GPU parity tests: one driver-like call sequence
replayed on the HIP library (through the C ABI) and on the CPU oracle."""
import numpy as np

import asl_amd
import oracle
from asl_amd import _lib

SLICE_KINDS = (_lib.BC_NOSLIP, _lib.BC_CONSTANT_PRESSURE, _lib.BC_CONSTANT_VELOCITY,
               _lib.BC_CONSTANT_PRESSURE_VELOCITY)


def pair(size, q=19, nu=0.1, dtype="f32", periodic=False, compressible=True):
    """(GPU lattice, oracle lattice) with the same parameters."""
    G = asl_amd.Lattice(size, q, dtype=dtype, periodic=periodic, nu=nu)
    O = oracle.OracleLattice(size, q, fp64=(dtype == "f64"), nu=nu, periodic=periodic)
    if not compressible:
        G.set_compressible(False)
        O.set_compressible(False)
    return G, O


def add_bc(G, O, kind, faces=0, mapv=None, p=0.0, u=None, rot=None, cdomain=None):
    if kind in SLICE_KINDS:
        return G.add_slice_bc(kind, faces, p=p, u=u), O.add_slice_bc(kind, faces, p=p, u=u)
    return (G.add_map_bc(kind, mapv, cdomain=cdomain, p=p, u=u, rot=rot),
            O.add_map_bc(kind, mapv, cdomain=cdomain, p=p, u=u, rot=rot))


def run(G, O, steps, bcs, pre_execute=True):
    """Driver loop: [execute BCs once], then steps x (LBGK execute; BCs)."""
    if pre_execute:
        for g, o in bcs:
            G.execute_bc(g)
            O.execute_bc(o)
        G.step(steps)
    elif steps:
        G.step(1)  # reads the initial ghost values raw
        for g, _ in bcs:
            G.execute_bc(g)
        G.step(steps - 1)
    order = [o for _, o in bcs]
    if steps:
        O.run(steps, order)


def interior_mask(G, extra=None):
    m = np.zeros(G.S, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    if extra is not None:
        m &= extra
    return m


def compare(G, O, mask=None, f=True):
    """Max abs differences of rho, u (and f) on mask (default: computation nodes)."""
    if mask is None:
        ct = G.cell_types()
        mask = (ct == _lib.CODE_FLUID) | (ct == _lib.CODE_FLUID_BC)
    out = {}
    rg, ro = G.rho().astype(np.float64), O.rho()
    ug, uo = G.velocity().astype(np.float64), O.velocity()
    out["rho"] = float(np.abs(rg[mask] - ro[mask]).max())
    out["u"] = float(max(np.abs(ug[d][mask] - uo[d][mask]).max() for d in range(3)))
    out["umax"] = float(max(np.abs(uo[d][mask]).max() for d in range(3)))
    if f:
        fg, fo = G.f().astype(np.float64), O.f()
        out["f"] = float(max(np.abs(fg[i][mask] - fo[i][mask]).max() for i in range(fg.shape[0])))
    out["cells"] = int(mask.sum())
    return out
