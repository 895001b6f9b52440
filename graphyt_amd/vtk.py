"""VTK XML writers for particle (.vtp, PolyData) and node (.vti, ImageData) output.

The reference's scripts write their results through an external `VTKWriter` module
(template_script.py:134-181):

    vtp = VTKWriter.VTPWriter(numParticles)
    vtp.AddPositions("Position", matpoints.getPosArr(), 3, VTKWriter.Order.ijk)
    vtp.AddArray("Stress", matpoints.getStressArr(), 6, VTKWriter.Order.ijk)
    vti = VTKWriter.VTIWriter([Nx, Ny, Nz], origin, cellsize)
    vti.AddArray("Velocity", nodes.getVelArr(), 3, VTKWriter.Order.ijk)
    ...
    vtp.Write("out.vtp"); vti.Write("out.vti")

That module is not part of the reference; this one has the same calls. Arrays are held by
reference and read when Write runs, as the scripts expect (they add the getter arrays
once and write after later steps). Arrays returned by the graphyt host module's getters
are views of its host buffers: before writing, each such owner is asked to fetch the
device state of the last step (its `_sync` method), so `iterate_GPU` loops need no extra
call. Data are written as raw appended binary (little-endian, UInt64 block headers),
which ParaView and VTK read.
"""
from __future__ import annotations

import enum
import os
import xml.sax.saxutils as su

import numpy as np

_VTK_TYPE = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
             np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
             np.dtype(np.uint8): "UInt8", np.dtype(np.int8): "Int8",
             np.dtype(np.uint32): "UInt32", np.dtype(np.uint64): "UInt64"}


class Order(enum.Enum):
    """Component order of a flat array: ijk = point-major ([p][component]), the layout of
    every graphyt array."""
    ijk = 0


def _sync_owner(arr) -> None:
    base = getattr(arr, "base", None)
    while base is not None:
        sync = getattr(base, "_sync", None)
        if callable(sync):
            sync()
            return
        base = getattr(base, "base", None)


class _Writer:
    def __init__(self, npoints: int):
        if npoints < 0:
            raise ValueError("number of points must be non-negative")
        self.npoints = int(npoints)
        self.arrays = []  # (name, array, ncomp)

    def _add(self, name: str, arr, ncomp: int, order=Order.ijk) -> None:
        if order is not Order.ijk and order != Order.ijk.value:
            raise ValueError("only Order.ijk is supported")
        if ncomp < 1:
            raise ValueError("ncomp must be >= 1")
        a = arr if isinstance(arr, np.ndarray) else np.asarray(arr)
        if a.size != self.npoints * ncomp:
            raise ValueError(f"{name}: {a.size} values for {self.npoints} points x {ncomp} components")
        self.arrays.append((str(name), a, int(ncomp)))

    def AddArray(self, name: str, arr, ncomp: int, order=Order.ijk) -> None:
        self._add(name, arr, ncomp, order)

    @staticmethod
    def _prep(a) -> np.ndarray:
        """Little-endian contiguous copy in a VTK type (others become Float64)."""
        a = np.asarray(a)
        dt = np.dtype(a.dtype).newbyteorder("=")
        if dt not in _VTK_TYPE:
            dt = np.dtype(np.float64)
        return np.ascontiguousarray(a, dtype=dt.newbyteorder("<"))

    def _write(self, path: str, head: str, sections: list, tail: str) -> None:
        """sections: list of (xml text before the array tags, [(tag attrs, array)])"""
        blocks, offset, parts = [], 0, [head]
        for text, arrays in sections:
            parts.append(text)
            for attrs, a in arrays:
                a = self._prep(a)
                raw = a.tobytes()
                blk = np.uint64(len(raw)).tobytes() + raw
                vtype = _VTK_TYPE[np.dtype(a.dtype).newbyteorder("=")]
                parts.append(f'        <DataArray type="{vtype}" {attrs} format="appended" offset="{offset}"/>\n')
                blocks.append(blk)
                offset += len(blk)
        parts.append(tail)
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        with open(path, "wb") as f:
            f.write("".join(parts).encode())
            f.write(b'  <AppendedData encoding="raw">\n   _')
            for b in blocks:
                f.write(b)
            f.write(b"\n  </AppendedData>\n</VTKFile>\n")

    def _synced(self):
        for _, a, _ in self.arrays:
            _sync_owner(a)


class VTPWriter(_Writer):
    """Particles as VTK PolyData: one vertex per particle, the positions as the points."""

    def __init__(self, numParticles: int):
        super().__init__(numParticles)
        self.positions = None  # (name, array)

    def AddPositions(self, name: str, arr, ncomp: int = 3, order=Order.ijk) -> None:
        if ncomp not in (2, 3):
            raise ValueError("positions have 2 or 3 components")
        a = arr if isinstance(arr, np.ndarray) else np.asarray(arr)
        if a.size != self.npoints * ncomp:
            raise ValueError(f"{name}: {a.size} values for {self.npoints} points x {ncomp} components")
        self.positions = (str(name), a, ncomp)

    def Write(self, filename: str) -> None:
        if self.positions is None:
            raise RuntimeError("AddPositions first")
        self._synced()
        _sync_owner(self.positions[1])
        n = self.npoints
        pname, parr, pcomp = self.positions
        pts = np.asarray(parr, np.float64).reshape(n, pcomp)
        if pcomp == 2:
            pts = np.concatenate([pts, np.zeros((n, 1))], axis=1)
        head = ('<?xml version="1.0"?>\n'
                '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
                '  <PolyData>\n'
                f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0" '
                'NumberOfStrips="0" NumberOfPolys="0">\n')
        pdata = [(f'Name={su.quoteattr(name)} NumberOfComponents="{nc}"', np.asarray(a).reshape(-1))
                 for name, a, nc in self.arrays]
        sections = [("      <PointData>\n", pdata),
                    ("      </PointData>\n      <Points>\n",
                     [(f'Name={su.quoteattr(pname)} NumberOfComponents="3"', pts.reshape(-1))]),
                    ("      </Points>\n      <Verts>\n",
                     [('Name="connectivity"', np.arange(n, dtype=np.int64)),
                      ('Name="offsets"', np.arange(1, n + 1, dtype=np.int64))])]
        tail = "      </Verts>\n    </Piece>\n  </PolyData>\n"
        self._write(filename, head, sections, tail)


class VTIWriter(_Writer):
    """Node fields as VTK ImageData on the uniform grid (node n = i + j Nx + k Nx Ny)."""

    def __init__(self, dims, origin, spacing):
        dims = [int(v) for v in dims]
        if len(dims) == 2:
            dims = dims + [1]
        if len(dims) != 3 or min(dims) < 1:
            raise ValueError("dims: 2 or 3 positive node counts")
        super().__init__(dims[0] * dims[1] * dims[2])
        self.dims = dims
        o = [float(v) for v in origin]
        self.origin = (o + [0.0, 0.0, 0.0])[:3]
        sp = [float(spacing)] * 3 if np.isscalar(spacing) else [float(v) for v in spacing]
        self.spacing = (sp + [1.0, 1.0, 1.0])[:3]

    def Write(self, filename: str) -> None:
        self._synced()
        nx, ny, nz = self.dims
        ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
        head = ('<?xml version="1.0"?>\n'
                '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
                f'  <ImageData WholeExtent="{ext}" Origin="{" ".join(repr(v) for v in self.origin)}" '
                f'Spacing="{" ".join(repr(v) for v in self.spacing)}">\n'
                f'    <Piece Extent="{ext}">\n')
        pdata = [(f'Name={su.quoteattr(name)} NumberOfComponents="{nc}"', np.asarray(a).reshape(-1))
                 for name, a, nc in self.arrays]
        sections = [("      <PointData>\n", pdata)]
        tail = "      </PointData>\n    </Piece>\n  </ImageData>\n"
        self._write(filename, head, sections, tail)


def read(filename: str) -> dict:
    """Arrays of a file written by these writers: {name: flat numpy array}, plus
    "_points" (PolyData) or "_extent", "_origin", "_spacing" (ImageData)."""
    import xml.etree.ElementTree as ET
    data = open(filename, "rb").read()
    start = data.index(b'<AppendedData encoding="raw">')
    us = data.index(b"_", start) + 1
    root = ET.fromstring(data[:start].decode() + "</VTKFile>")
    types = {v: k for k, v in _VTK_TYPE.items()}
    out = {}
    for el in root.iter("DataArray"):
        off = us + int(el.get("offset"))
        nbytes = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        arr = np.frombuffer(data[off + 8:off + 8 + nbytes], types[el.get("type")]).copy()
        out[el.get("Name")] = arr
    img = root.find("ImageData")
    if img is not None:
        out["_extent"] = [int(v) for v in img.get("WholeExtent").split()]
        out["_origin"] = [float(v) for v in img.get("Origin").split()]
        out["_spacing"] = [float(v) for v in img.get("Spacing").split()]
    return out
