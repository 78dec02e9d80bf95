"""Minimal VTK XML writers (ImageData .vti, PolyData .vtp) with raw appended binary arrays.

Each array is written once, little-endian, after a UInt64 byte count (``header_type="UInt64"``,
``encoding="raw"``), the layout ParaView, VTK and pyvista read. Cell arrays of an ImageData are
in VTK order: x fastest, then y, then z.
"""

from __future__ import annotations

import pathlib

import numpy as np

_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
              np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8",
              np.dtype(np.int64): "Int64"}


def _vtk_type(a: np.ndarray) -> str:
    try:
        return _VTK_TYPES[a.dtype]
    except KeyError as exc:
        raise TypeError(f"no VTK type for {a.dtype}") from exc


class _Appended:
    def __init__(self):
        self.blobs: list[bytes] = []
        self.offset = 0

    def add(self, a: np.ndarray) -> int:
        raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
        off = self.offset
        blob = np.uint64(len(raw)).tobytes() + raw
        self.blobs.append(blob)
        self.offset += len(blob)
        return off


def _data_array(app: _Appended, name: str, a: np.ndarray, ncomp: int = 1) -> str:
    off = app.add(a)
    comp = f' NumberOfComponents="{ncomp}"' if ncomp != 1 else ""
    return (f'        <DataArray type="{_vtk_type(a)}" Name="{name}"{comp} format="appended" '
            f'offset="{off}"/>')


def _write(path: pathlib.Path, head: list[str], app: _Appended, tail: list[str]) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(("\n".join(head + tail) + "\n").encode())
        fh.write(b'  <AppendedData encoding="raw">\n   _')
        for blob in app.blobs:
            fh.write(blob)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def write_image_data(path, n_cells: tuple[int, int, int], spacing: float,
                     origin: tuple[float, float, float], cell_data: dict[str, np.ndarray]) -> None:
    """ImageData of nx*ny*nz cells; every array holds nx*ny*nz values in VTK order."""
    nx, ny, nz = n_cells
    n = nx * ny * nz
    app = _Appended()
    arrays = []
    for name, a in cell_data.items():
        a = np.asarray(a).reshape(-1)
        if a.size != n:
            raise ValueError(f"{name}: {a.size} values for {n} cells")
        arrays.append(_data_array(app, name, a))
    ext = f"0 {nx} 0 {ny} 0 {nz}"
    head = ['<?xml version="1.0"?>',
            '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64">',
            f'  <ImageData WholeExtent="{ext}" Origin="{origin[0]!r} {origin[1]!r} '
            f'{origin[2]!r}" Spacing="{spacing!r} {spacing!r} {spacing!r}">',
            f'    <Piece Extent="{ext}">',
            "      <PointData>", "      </PointData>", "      <CellData>", *arrays,
            "      </CellData>", "    </Piece>", "  </ImageData>"]
    _write(pathlib.Path(path), head, app, [])


def write_poly_data(path, points: np.ndarray, point_data: dict[str, np.ndarray],
                    verts: bool = False, triangles: np.ndarray | None = None) -> None:
    """PolyData of ``points`` with either one vertex cell per point (``verts``) or the
    triangles (m, 3) of point indices."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    n = pts.shape[0]
    app = _Appended()
    parts = ["      <Points>", _data_array(app, "Points", pts, 3), "      </Points>",
             "      <PointData>"]
    for name, a in point_data.items():
        a = np.asarray(a).reshape(-1)
        if a.size != n:
            raise ValueError(f"{name}: {a.size} values for {n} points")
        parts.append(_data_array(app, name, a))
    parts.append("      </PointData>")
    n_verts = n_polys = 0
    if verts:
        n_verts = n
        parts += ["      <Verts>",
                  _data_array(app, "connectivity", np.arange(n, dtype=np.int64)),
                  _data_array(app, "offsets", np.arange(1, n + 1, dtype=np.int64)),
                  "      </Verts>"]
    if triangles is not None:
        tri = np.ascontiguousarray(triangles, dtype=np.int64).reshape(-1, 3)
        n_polys = tri.shape[0]
        parts += ["      <Polys>",
                  _data_array(app, "connectivity", tri.reshape(-1)),
                  _data_array(app, "offsets", np.arange(3, 3 * n_polys + 1, 3, dtype=np.int64)),
                  "      </Polys>"]
    head = ['<?xml version="1.0"?>',
            '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64">',
            "  <PolyData>",
            f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n_verts}" NumberOfLines="0" '
            f'NumberOfStrips="0" NumberOfPolys="{n_polys}">',
            *parts, "    </Piece>", "  </PolyData>"]
    _write(pathlib.Path(path), head, app, [])


def write_poly_points(path, points: np.ndarray, point_data: dict[str, np.ndarray]) -> None:
    """PolyData of vertices: one vertex cell per point."""
    write_poly_data(path, points, point_data, verts=True)


def read_vtk_arrays(path) -> tuple[dict, dict[str, np.ndarray]]:
    """Read back a file written here: (attributes, {name: array}); the attributes are the
    ImageData extent, origin and spacing, or the PolyData piece counts. Arrays of a PolyData's
    Verts and Polys come back as "<Section>/connectivity" and "<Section>/offsets"."""
    import re

    data = pathlib.Path(path).read_bytes()
    head, _, rest = data.partition(b'<AppendedData encoding="raw">')
    start = rest.index(b"_") + 1
    blob = rest[start:]
    text = head.decode()
    m = re.search(r'<ImageData WholeExtent="([^"]+)" Origin="([^"]+)" Spacing="([^"]+)"', text)
    if m:
        attrs = {"extent": [int(v) for v in m.group(1).split()],
                 "origin": [float(v) for v in m.group(2).split()],
                 "spacing": [float(v) for v in m.group(3).split()]}
    else:
        attrs = {k: int(v) for k, v in re.findall(r'NumberOf(\w+)="(\d+)"', text)}
    inv = {v: k for k, v in _VTK_TYPES.items()}
    out = {}
    section = ""
    for line in text.splitlines():
        sec = re.match(r"\s*<(Verts|Polys)>", line)
        if sec:
            section = sec.group(1) + "/"
            continue
        if re.match(r"\s*</(Verts|Polys)>", line):
            section = ""
            continue
        da = re.search(r'<DataArray type="(\w+)" Name="([^"]+)"(?: NumberOfComponents="\d+")? '
                       r'format="appended" offset="(\d+)"/>', line)
        if not da:
            continue
        t, name, off = da.groups()
        name = section + name if name in ("connectivity", "offsets") else name
        off = int(off)
        nbytes = int(np.frombuffer(blob[off:off + 8], dtype="<u8")[0])
        out[name] = np.frombuffer(blob[off + 8:off + 8 + nbytes], dtype=inv[t].newbyteorder("<"))
    return attrs, out


def read_image_cell_data(path) -> tuple[dict, dict[str, np.ndarray]]:
    return read_vtk_arrays(path)
