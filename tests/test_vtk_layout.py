"""The particle .vtu / .pvd layout pinned byte for byte against the VTK XML file format
(VTK's "VTK File Formats" documentation: an UnstructuredGrid piece with Points, Cells
(connectivity, offsets, types) and PointData; binary DataArrays written inline as
base64 of a header holding the payload's byte count -- an unsigned 64-bit little-endian
integer with header_type="UInt64" -- followed by the little-endian payload; cell type 1
is VTK_VERTEX).  The expected text of every array is assembled here element by element
with struct, not with the writer's numpy path."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from scaletrack import vtk


def b64(fmt, values):
    payload = b"".join(struct.pack("<" + fmt, v) for v in values)
    return base64.b64encode(struct.pack("<Q", len(payload)) + payload).decode("ascii")


@pytest.mark.parametrize("dtype,fmt,name", [(np.float64, "d", "Float64"), (np.float32, "f", "Float32")])
def test_vtu_bytes(tmp_path, dtype, fmt, name):
    X = np.array([0.25, 0.5], dtype)
    Y = np.array([0.125, 0.75], dtype)
    Z = np.array([1.0, 0.0625], dtype)
    U = np.array([1.5, -2.0], dtype)
    V = np.array([0.0, 3.25], dtype)
    W = np.array([-0.5, 0.25], dtype)
    d = np.array([3e-5, 5e-5], dtype)
    path = vtk.write_vtu(str(tmp_path / "p.vtu"), X, Y, Z, U, V, W, d)
    text = open(path, "rb").read()
    assert text.startswith(b"<?xml version='1.0' encoding='utf-8'?>")
    root = ET.fromstring(text)
    assert root.tag == "VTKFile"
    assert root.attrib == {"type": "UnstructuredGrid", "version": "1.0", "byte_order": "LittleEndian",
                           "header_type": "UInt64"}
    (grid,) = list(root)
    assert grid.tag == "UnstructuredGrid"
    (piece,) = list(grid)
    assert piece.attrib == {"NumberOfPoints": "2", "NumberOfCells": "2"}
    assert [c.tag for c in piece] == ["Points", "Cells", "PointData"]
    arrays = [(parent.tag, el) for parent in piece for el in parent]
    expect = [
        ("Points", "Points", name, "3", b64(fmt, [X[0], Y[0], Z[0], X[1], Y[1], Z[1]])),
        ("Cells", "connectivity", "Int64", "1", b64("q", [0, 1])),
        ("Cells", "offsets", "Int64", "1", b64("q", [1, 2])),
        ("Cells", "types", "UInt8", "1", b64("B", [1, 1])),
        ("PointData", "U", name, "3", b64(fmt, [U[0], V[0], W[0], U[1], V[1], W[1]])),
        ("PointData", "d", name, "1", b64(fmt, list(d))),
    ]
    assert len(arrays) == len(expect)
    for (ptag, el), (etag, ename, etype, ncomp, etext) in zip(arrays, expect):
        assert ptag == etag and el.tag == "DataArray"
        assert el.attrib == {"type": etype, "Name": ename, "format": "binary", "NumberOfComponents": ncomp}
        assert el.text == etext, ename


def test_pvd_collection(tmp_path):
    coll = vtk.ParaviewCollection(str(tmp_path / "particleTimeSeries.pvd"))
    (tmp_path / "dataVTK").mkdir()
    coll.add(0.005, str(tmp_path / "dataVTK" / "particleFields_0.005.vtu"))
    coll.add(0.01, str(tmp_path / "dataVTK" / "particleFields_0.01.vtu"))
    root = ET.parse(coll.save()).getroot()
    assert root.attrib == {"type": "Collection", "version": "1.0", "byte_order": "LittleEndian"}
    (c,) = list(root)
    assert c.tag == "Collection"
    assert [d.attrib for d in c] == [
        {"timestep": "0.005", "part": "0", "file": "dataVTK/particleFields_0.005.vtu"},
        {"timestep": "0.01", "part": "0", "file": "dataVTK/particleFields_0.01.vtu"}]
