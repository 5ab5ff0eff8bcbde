"""Reader for the VTK XML files of the VTK output (StructuredGrid with raw
appended data, UInt64 headers, and the multiblock index): enough to check
their content in the tests without the VTK library."""
import re
import xml.etree.ElementTree as ET

import numpy as np


def read_vts(path):
    raw = open(path, "rb").read()
    mark = raw.index(b"<AppendedData")
    start = raw.index(b"_", mark) + 1
    head = raw[:mark].decode() + "</VTKFile>"
    root = ET.fromstring(re.sub(r"<AppendedData.*", "", head, flags=re.S))
    piece = root.find("StructuredGrid/Piece")
    ext = [int(v) for v in piece.get("Extent").split()]
    shape = (ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4])   # cells along i, j, k

    def block(off):
        n = int(np.frombuffer(raw, dtype="<u8", count=1, offset=start + off)[0])
        return np.frombuffer(raw, dtype="<f4", count=n // 4, offset=start + off + 8)

    cells = {}
    for da in piece.find("CellData"):
        assert da.get("type") == "Float32" and da.get("format") == "appended"
        a = block(int(da.get("offset")))
        nc = int(da.get("NumberOfComponents", "1"))
        # VTK order: i fastest, then j, then k
        a = a.reshape(shape[2], shape[1], shape[0], nc) if nc > 1 else a.reshape(shape[2], shape[1], shape[0])
        cells[da.get("Name")] = a
    pda = piece.find("Points/DataArray")
    pts = block(int(pda.get("offset"))).reshape(shape[2] + 1, shape[1] + 1, shape[0] + 1, 3)
    return dict(root=root, extent=ext, shape=shape, cells=cells, points=pts)


def read_vtm(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "vtkMultiBlockDataSet"
    return [(int(d.get("index")), d.get("file")) for d in root.find("vtkMultiBlockDataSet")]
