"""Reader of the .vti files the compat layer's WriterVTKXML writes (VTK XML
ImageData, raw appended data, UInt64 block headers, Float32 arrays, x index
fastest): arrays returned in ASL's index order (x, y, z) over the internal
cells, vectors as (3, nx, ny, nz)."""
import re

import numpy as np


def read_vti(path):
    data = open(path, "rb").read()
    head_end = data.index(b"<AppendedData encoding=\"raw\">")
    head = data[:head_end].decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    n = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    arrays = re.findall(r'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)" '
                        r'format="appended" offset="(\d+)"/>', head)
    base = data.index(b"_", head_end) + 1
    out = {}
    for name, nc, off in arrays:
        nc, off = int(nc), int(off)
        nbytes = int(np.frombuffer(data, dtype="<u8", count=1, offset=base + off)[0])
        a = np.frombuffer(data, dtype="<f4", count=nbytes // 4, offset=base + off + 8)
        a = a.reshape(n[2], n[1], n[0], nc).transpose(3, 2, 1, 0)
        out[name] = a[0] if nc == 1 else a
    return out
