"""Particle output in VTK XML, as the reference writes it with WriteVTK
(run/icoJuliaParcelFoam/cavity3D/asyncParallelTracking.jl:1367-1430): one .vtu per
write with the parcel positions as points, one VTK_VERTEX cell per parcel, point data
"U" (velocity, 3 components) and "d" (diameter), collected in a .pvd time series.
Arrays are written inline as base64 binary with a UInt64 byte-count header.
"""
import base64
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

VTK_VERTEX = 1
_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64",
          np.dtype(np.uint8): "UInt8"}


def _data_array(parent, name, arr, ncomp=1):
    arr = np.ascontiguousarray(arr)
    el = ET.SubElement(parent, "DataArray", type=_TYPES[arr.dtype], Name=name, format="binary",
                       NumberOfComponents=str(ncomp))
    raw = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
    el.text = base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")
    return el


def write_vtu(path, X, Y, Z, U, V, W, d):
    """One particle snapshot; returns the path written."""
    n = len(X)
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(n), NumberOfCells=str(n))
    pts = ET.SubElement(piece, "Points")
    _data_array(pts, "Points", np.stack([X, Y, Z], axis=1), 3)
    cells = ET.SubElement(piece, "Cells")
    _data_array(cells, "connectivity", np.arange(n, dtype=np.int64))
    _data_array(cells, "offsets", np.arange(1, n + 1, dtype=np.int64))
    _data_array(cells, "types", np.full(n, VTK_VERTEX, dtype=np.uint8))
    pd = ET.SubElement(piece, "PointData")
    _data_array(pd, "U", np.stack([U, V, W], axis=1), 3)
    _data_array(pd, "d", np.asarray(d))
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
    return path


class ParaviewCollection:
    """The .pvd time series of the written snapshots (paraview_collection / vtk_save)."""

    def __init__(self, path="particleTimeSeries.pvd"):
        self.path = path
        self.entries = []

    def add(self, t, vtu_path):
        self.entries.append((float(t), vtu_path))

    def save(self):
        root = ET.Element("VTKFile", type="Collection", version="1.0", byte_order="LittleEndian")
        coll = ET.SubElement(root, "Collection")
        base = os.path.dirname(os.path.abspath(self.path))
        for t, f in self.entries:
            ET.SubElement(coll, "DataSet", timestep=repr(t), part="0",
                          file=os.path.relpath(os.path.abspath(f), base))
        ET.ElementTree(root).write(self.path, xml_declaration=True, encoding="utf-8")
        return self.path


def write_chunk(chunk, t, collection=None, directory="dataVTK"):
    """write(chunk, ::CPU) of :1367-1390: dataVTK/particleFields_<t>.vtu with t rounded
    to 4 significant digits, registered in the collection."""
    os.makedirs(directory, exist_ok=True)
    tr = float(f"{t:.4g}")
    path = os.path.join(directory, f"particleFields_{tr}.vtu")
    write_vtu(path, chunk.X, chunk.Y, chunk.Z, chunk.U, chunk.V, chunk.W, chunk.d)
    if collection is not None:
        collection.add(tr, path)
    return path


def read_vtu(path):
    """Arrays of a file written by write_vtu (for checks)."""
    root = ET.parse(path).getroot()
    out = {}
    for el in root.iter("DataArray"):
        raw = base64.b64decode(el.text)
        nbytes = struct.unpack("<Q", raw[:8])[0]
        dt = {v: k for k, v in _TYPES.items()}[el.get("type")]
        a = np.frombuffer(raw[8:8 + nbytes], dtype=dt.newbyteorder("<"))
        ncomp = int(el.get("NumberOfComponents", "1"))
        out[el.get("Name")] = a.reshape(-1, ncomp) if ncomp > 1 else a
    return out
