"""ASL Binary Dump (ABD) files: the stream of ``src/writers/aslABDFormat.h:36-135``
(native byte order, no tags; a vector or string is its unsigned 32-bit length and its
elements; a Block is dx, size, position) in the layout ``asl::writeABD`` writes here:
name, block, number of components, then every component as doubles over the
ghost-inclusive block (ASL's layout, last index fastest).

``Lattice.save_abd`` / ``Lattice.load_abd`` use it for population checkpoints.
"""
import struct

import numpy as np


def _vec(fmt, values):
    values = list(values)
    return struct.pack("=I", len(values)) + struct.pack("=%d%s" % (len(values), fmt), *values)


def write_abd(path, name, components, dx=1.0, position=None):
    """Write components (array of shape (n,) + S) with the block (dx, S, position)."""
    comp = np.ascontiguousarray(np.asarray(components, dtype=np.float64))
    S = comp.shape[1:]
    if position is None:
        position = (-dx,) * len(S)
    raw = name.encode()
    with open(path, "wb") as f:
        f.write(struct.pack("=I", len(raw)) + raw)
        f.write(struct.pack("=d", float(dx)) + _vec("i", S) + _vec("d", position))
        f.write(struct.pack("=I", comp.shape[0]))
        f.write(comp.tobytes())


def read_abd(path):
    """(name, dx, size, position, components of shape (n,) + size)."""
    with open(path, "rb") as f:
        data = f.read()
    at = 0

    def take(fmt):
        nonlocal at
        v = struct.unpack_from(fmt, data, at)
        at += struct.calcsize(fmt)
        return v

    (n,) = take("=I")
    name = data[at:at + n].decode()
    at += n
    (dx,) = take("=d")
    (ns,) = take("=I")
    size = take("=%di" % ns)
    (npos,) = take("=I")
    position = take("=%dd" % npos)
    (nc,) = take("=I")
    count = nc * int(np.prod(size))
    if len(data) - at != 8 * count:
        raise ValueError("ABD file %s: %d bytes of data, %d expected" % (path, len(data) - at, 8 * count))
    comp = np.frombuffer(data, dtype=np.float64, count=count, offset=at).reshape((nc,) + tuple(size))
    return name, dx, tuple(size), tuple(position), comp
