"""Level-set (distance-function) maps on ASL's ghost-inclusive block, host side.

Restates the driver-side map construction ``initData(map, DF)`` /
``generateDataContainer_SP(block, DF, 1, type)``
(``src/aslGeomInc.cxx:187-221``): every cell of the ghost-inclusive block
gets the distance function at its position, evaluated in float32 like
ASL's generated kernel.

* position of ghost-inclusive index I along axis d:
  ``float(position_d - dx) + float(dx) * float(I_d)``
  (``Index2PositionACL``, ``src/math/aslIndex2Position.cxx:64-108``;
  ``offset(b, 1)``, ``src/data/aslBlocks.cxx:28-37``)
* sphere: ``sqrt(|pos - c|^2) - r`` (``DistFSphere``,
  ``src/math/aslDistanceFunction.cxx:98-101``)
* normalisation: ``clamp(d / (1.8 dx), -1, 1)`` with the factor rounded to
  float (``DistFNormalization``, ``src/math/aslDistanceFunction.cxx:142-164``)
* CSG: union = min, intersection = max, complement = negation
  (``:64-89``)

The maps are inputs of the map boundary conditions; they are passed to the
library as float arrays in ASL's layout.
"""
import numpy as np

F = np.float32


def positions(size, dx=1.0, position=(0.0, 0.0, 0.0)):
    """Cell-centre positions (3 float32 arrays of shape size+2) of the ghost-inclusive block."""
    S = tuple(int(s) + 2 for s in size)
    out = []
    for d in range(3):
        i = np.arange(S[d], dtype=F)
        p = F(position[d] - dx) + F(dx) * i
        shape = [1, 1, 1]
        shape[d] = S[d]
        out.append(np.broadcast_to(p.reshape(shape), S))
    return out


def sphere(size, radius, centre, dx=1.0, position=(0.0, 0.0, 0.0)):
    """generateDFSphere(r, c) sampled on the ghost-inclusive block (float32)."""
    x, y, z = positions(size, dx, position)
    d0 = x - F(centre[0])
    d1 = y - F(centre[1])
    d2 = z - F(centre[2])
    l2 = (d0 * d0 + d1 * d1) + d2 * d2
    return (np.sqrt(l2) - F(radius)).astype(F)


def cylinder_inf(size, radius, axis, centre, dx=1.0, position=(0.0, 0.0, 0.0)):
    """generateDFCylinderInf(r, axis, c): distance to an infinite cylinder (float32).

    ``sqrt(l2((pos-c) - ((pos-c).a) a)) - r`` with a = axis/|axis|
    (``DistFCylinder``, ``src/math/aslDistanceFunction.cxx:103-116``).
    """
    a = np.asarray(axis, dtype=np.float64)
    a = (a / np.sqrt((a * a).sum())).astype(F)
    p = positions(size, dx, position)
    d = [p[i] - F(centre[i]) for i in range(3)]
    proj = (d[0] * a[0] + d[1] * a[1]) + d[2] * a[2]
    r = [d[i] - proj * a[i] for i in range(3)]
    l2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
    return (np.sqrt(l2) - F(radius)).astype(F)


def normalize(df, dx=1.0):
    """normalize(df, dx): clamp(df * float(1/dx/1.8), -1, 1)."""
    dist = F(1.0 / dx / 1.8) * np.asarray(df, dtype=F)
    top = np.where(dist > F(1.0), F(1.0), dist)
    return np.where(dist < F(-1.0), F(-1.0), top).astype(F)


def union(a, b):
    return np.minimum(a, b)


def intersection(a, b):
    return np.maximum(a, b)


def complement(a):
    return -a
