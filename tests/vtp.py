"""A small VTK XML PolyData writer for the tests, following the file format
VTK's vtkXMLPolyDataWriter produces: DataArrays in ascii, binary (base64
inline) or appended (raw or base64) form, UInt32/UInt64 headers, optional
vtkZLibDataCompressor blocks, either byte order.  Binary data is
[header][bytes]; the header is the byte count, or (compressed)
[nblocks, block size, last block size, compressed size of each block]; in
base64 the header and the data are encoded as separate units."""
import base64
import zlib

import numpy as np

_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32",
          np.dtype(np.int64): "Int64", np.dtype(np.uint32): "UInt32", np.dtype(np.uint64): "UInt64"}


def _units(arr, header, compress, big, block):
    """(header bytes, data bytes) of one binary DataArray."""
    e = ">" if big else "<"
    data = np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder(e)).tobytes()
    ht = np.dtype(e + ("u4" if header == "UInt32" else "u8"))
    if not compress:
        return np.array([len(data)], dtype=ht).tobytes(), data
    blocks = [data[i:i + block] for i in range(0, len(data), block)] or [b""]
    comp = [zlib.compress(b) for b in blocks]
    last = len(blocks[-1]) if len(blocks[-1]) != block else 0
    hdr = np.array([len(blocks), block, last] + [len(c) for c in comp], dtype=ht).tobytes()
    return hdr, b"".join(comp)


def write_vtp(path, pieces, form="ascii", header="UInt32", compress=False, big=False, appended_encoding="raw",
              block=64):
    """pieces: list of dicts with "points" (n, 3) float32/float64 and optional
    "polys" / "strips" (lists of point-index lists) and "verts" / "lines"
    (written, never surfaces); connectivity written as int64."""
    app = []  # appended units
    app_len = 0

    def array(arr, name=None, ncomp=None):
        nonlocal app_len
        attrs = 'type="%s"' % _TYPES[arr.dtype]
        if name:
            attrs += ' Name="%s"' % name
        if ncomp:
            attrs += ' NumberOfComponents="%d"' % ncomp
        if form == "ascii":
            fmt = "%.9g" if arr.dtype == np.float32 else ("%.17g" if arr.dtype.kind == "f" else "%d")
            return '<DataArray %s format="ascii">\n%s\n</DataArray>\n' % (
                attrs, " ".join(fmt % v for v in arr.ravel()))
        hdr, data = _units(arr, header, compress, big, block)
        if form == "binary":
            text = base64.b64encode(hdr).decode() + base64.b64encode(data).decode()
            return '<DataArray %s format="binary">\n%s\n</DataArray>\n' % (attrs, text)
        if appended_encoding == "raw":
            unit = hdr + data
        else:
            unit = base64.b64encode(hdr) + base64.b64encode(data)
        s = '<DataArray %s format="appended" offset="%d"/>\n' % (attrs, app_len)
        app.append(unit)
        app_len += len(unit)
        return s

    def cells(lists):
        con = np.array([i for c in lists for i in c], dtype=np.int64)
        off = np.cumsum([len(c) for c in lists]).astype(np.int64)
        return array(con, "connectivity") + array(off, "offsets")

    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PolyData" version="1.0" byte_order="%s" header_type="%s"%s>\n'
           % ("BigEndian" if big else "LittleEndian", header,
              ' compressor="vtkZLibDataCompressor"' if compress else ""),
           "<PolyData>\n"]
    for p in pieces:
        pts = np.asarray(p["points"])
        kinds = ("verts", "lines", "strips", "polys")
        counts = " ".join('NumberOf%s="%d"' % (k.capitalize(), len(p.get(k, []))) for k in kinds)
        out.append('<Piece NumberOfPoints="%d" %s>\n' % (len(pts), counts))
        out.append("<PointData>\n%s</PointData>\n" % array(np.arange(len(pts), dtype=np.int32), "id"))
        out.append("<Points>\n%s</Points>\n" % array(pts.reshape(-1), "Points", 3))
        for k in kinds:
            if p.get(k):
                out.append("<%s>\n%s</%s>\n" % (k.capitalize(), cells(p[k]), k.capitalize()))
        out.append("</Piece>\n")
    out.append("</PolyData>\n")
    body = "".join(out).encode()
    if app:
        body += ('<AppendedData encoding="%s">\n_' % appended_encoding).encode() + b"".join(app)
        body += b"\n</AppendedData>\n"
    body += b"</VTKFile>\n"
    with open(path, "wb") as f:
        f.write(body)


def triangles(pieces):
    """The triangles readSurface takes from the pieces: polygons as fans from
    their first point, strips as consecutive triples (float64)."""
    tri = []
    for p in pieces:
        pts = np.asarray(p["points"], dtype=np.float64)
        for poly in p.get("polys", []):
            for k in range(1, len(poly) - 1):
                tri.append(pts[[poly[0], poly[k], poly[k + 1]]])
        for st in p.get("strips", []):
            for k in range(len(st) - 2):
                tri.append(pts[[st[k], st[k + 1], st[k + 2]]])
    return np.array(tri, dtype=np.float64).reshape(-1, 3, 3)
