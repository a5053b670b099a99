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


# --- triangulated surfaces (asl::readSurface) ------------------------------

def read_stl(path):
    """Triangles of a binary or ASCII STL file, shape (n, 3, 3) float64 in the
    file's x y z (values rounded to float, as VTK's reader keeps them):
    the reading half of ``readSurface`` (``src/readers/aslVTKFormatReaders.cxx:266-305``)."""
    import ctypes

    from . import _lib
    p = str(path).encode()
    n = ctypes.c_longlong(0)
    _lib.call("aslhip_stl_count", p, ctypes.byref(n))
    tri = np.empty((n.value, 3, 3), dtype=np.float64)
    if n.value:
        _lib.call("aslhip_stl_read", p, tri.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value)
    return tri


def read_vtp(path):
    """Triangles of a VTK XML PolyData file (polygons as fans from their first
    point, strips as consecutive triples), shape (n, 3, 3) float64 in the
    file's x y z: the reading half of ``readSurface`` for ``.vtp`` files
    (``src/readers/aslVTKFormatReaders.cxx:219-265``)."""
    import ctypes

    from . import _lib
    p = str(path).encode()
    n = ctypes.c_longlong(0)
    _lib.call("aslhip_vtp_count", p, ctypes.byref(n))
    tri = np.empty((n.value, 3, 3), dtype=np.float64)
    if n.value:
        _lib.call("aslhip_vtp_read", p, tri.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value)
    return tri


def read_surface_file(path):
    """Triangles of an ``.stl`` or ``.vtp`` surface, by extension (as readSurface dispatches)."""
    ext = str(path).rsplit(".", 1)[-1].lower() if "." in str(path) else ""
    if ext == "stl":
        return read_stl(path)
    if ext == "vtp":
        return read_vtp(path)
    raise ValueError("readSurface: file format not supported: %s" % path)


def surface_map(tri, shape, pos, h, cap, shift, div1, div2):
    """Distance map of triangles (n, 3, 3) in ASL axis order on the nodes
    ``pos + h * index`` of a grid of ``shape`` (on the GPU; see
    ``aslhip_surface_distance_map``): float32 array of that shape."""
    import ctypes

    from . import _lib
    t = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(-1))
    sz = np.ascontiguousarray(np.asarray(shape, dtype=np.int32).reshape(3))
    pa = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    ha = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (3,)))
    out = np.empty(tuple(int(s) for s in sz), dtype=np.float32)
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.call("aslhip_surface_distance_map", t.ctypes.data_as(dp), t.size // 9,
              sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), pa.ctypes.data_as(dp), ha.ctypes.data_as(dp),
              float(cap), float(shift), float(div1), float(div2), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def stl_to_asl_axes(tri):
    """readSurface's grid takes the file's x as ASL's last axis: (x, y, z) -> (z, y, x)."""
    return np.asarray(tri)[..., ::-1]


def block_nodes(size, dx, position):
    """Node origin and spacing per axis of readSurface(file, Block) on the
    ghost-inclusive grid of ``size`` nodes at ``position``: VTK's model bounds
    [position, position + (size - 1) dx] over size - 1 intervals
    (``src/readers/aslVTKFormatReaders.cxx:155-183``)."""
    pos = np.asarray(position, dtype=np.float64)
    s = np.asarray(size, dtype=np.float64)
    top = pos + (s - 1.0) * dx
    return pos, (top - pos) / (s - 1.0)


def read_surface(path, size, dx, position=(0.0, 0.0, 0.0)):
    """``readSurface(file, Block b)``: the map (float32, shape ``size``, the
    ghost-inclusive grid of the lattice whose internal block is ``size - 2``
    at ``position + dx``) of the STL surface:
    ``(d - 1.2 dx) / dx / 1.8`` with d the distance capped at 3.6 dx."""
    tri = stl_to_asl_axes(read_stl(path))
    pos, h = block_nodes(size, dx, position)
    return surface_map(tri, size, pos, h, 3.6 * dx, 1.2 * dx, dx, 1.8)


# --- distance-function expressions evaluated on the GPU (initData(map, DF)) --
# Postfix programs of the ops in include/aslhip.h (ASLHIP_DF_*), built like
# the generators of src/aslGeomInc.cxx:33-221; df_map samples one on a grid.

def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.sqrt((v * v).sum())


def df_sphere(r, c):
    return [1.0, float(c[0]), float(c[1]), float(c[2]), float(r)]


def df_cylinder_inf(r, axis, c):
    o = _unit(axis)
    return [2.0, float(c[0]), float(c[1]), float(c[2]), o[0], o[1], o[2], float(r)]


def df_plane(n, p0):
    """DistFPlane(n, p0): p . n/|n| - (p0 . n)/|n| (src/math/aslDistanceFunction.cxx:133-143)."""
    n = np.asarray(n, dtype=np.float64)
    o = _unit(n)
    return [4.0, o[0], o[1], o[2], float(np.dot(np.asarray(p0, dtype=np.float64), n) / np.sqrt((n * n).sum()))]


def df_and(a, b):
    return list(a) + list(b) + [5.0]


def df_or(a, b):
    return list(a) + list(b) + [6.0]


def df_neg(a):
    return list(a) + [7.0]


def df_normalize(a, dx=1.0):
    return list(a) + [8.0, 1.0 / dx / 1.8]


def df_cylinder(r, l, c):
    """generateDFCylinder: the infinite cylinder cut by the planes at c +- l/2."""
    l, c = np.asarray(l, dtype=np.float64), np.asarray(c, dtype=np.float64)
    return df_and(df_and(df_cylinder_inf(r, l, c), df_plane(l, c + 0.5 * l)), df_plane(-l, c - 0.5 * l))


def df_cone(r, l, a):
    """generateDFCone(r, l, a): DistFCone(r/|l|, l, a) cut at the apex and the base."""
    l, a = np.asarray(l, dtype=np.float64), np.asarray(a, dtype=np.float64)
    o = _unit(l)
    cone = [3.0, a[0], a[1], a[2], o[0], o[1], o[2], float(np.tan(r / np.sqrt((l * l).sum())))]
    return df_and(df_and(cone, df_plane(l, a + l)), df_plane(-l, a))


def df_map(prog, shape, pos, dx, fp64=False):
    """The map of a program on the nodes pos + dx * index of a grid of ``shape``
    (on the GPU; see ``aslhip_df_map``), float32 (or float64) of that shape."""
    import ctypes

    from . import _lib
    pr = np.ascontiguousarray(np.asarray(prog, dtype=np.float64).reshape(-1))
    sz = np.ascontiguousarray(np.asarray(shape, dtype=np.int32).reshape(3))
    pa = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    out = np.empty(tuple(int(s) for s in sz), dtype=np.float64 if fp64 else np.float32)
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.call("aslhip_df_map", pr.ctypes.data_as(dp), pr.size, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
              pa.ctypes.data_as(dp), float(dx), int(bool(fp64)), out.ctypes.data)
    return out
