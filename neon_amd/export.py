"""On-disk exports of the moment fields, in the reference's formats.

* Legacy VTK (``write_vtk`` / ``read_vtk``): the files Field::ioToVtk writes
  (libNeonDomain/include/Neon/domain/interface/FieldBase_imp.h:297-318 over
  libNeonCore/include/Neon/core/tools/io/ioToVTK.h:229-445): a
  STRUCTURED_POINTS header, CELL_DATA then POINT_DATA, one SCALARS block per
  field ending in ``METADATA`` / ``INFORMATION 0``; ASCII values printed like
  a default C++ stream (``%g``, six significant digits) with a space after
  each value and a newline per x row, BINARY values big-endian followed by a
  newline.  The C++ writer in include/neon_amd.hpp (Field::ioToVtk) produces
  the same bytes.
* XML ImageData (``write_vti`` / ``read_vti``): the files Neon::ioToVTI
  writes (libNeonCore/include/Neon/core/tools/io/ioToVti.h:663-719): one
  Float64 DataArray per field, PointData then CellData; ASCII values in
  fixed notation with 17 decimals, each followed by a space, a tab after
  each voxel, a newline after each x row and each z plane; BINARY arrays as
  little-endian binary64 behind a UInt32 byte count, base64 in one stream.
  The C++ writer (neon_amd::ioToVTI in include/neon_amd.hpp) produces the
  same bytes.
* Centre lines (``centerlines`` / ``write_dat``): the two
  NeonUniformLBM_<iter>_{X,Y}.dat files of the cavity benchmark
  (benchmarks/lbm-lid-driven-cavity-flow/src/RunCavityTwoPop.cu:113-149).
"""
from __future__ import annotations

import base64
import xml.etree.ElementTree as ET

import numpy as np

_TYPES = {np.dtype(np.float64): "double", np.dtype(np.float32): "float",
          np.dtype(np.int32): "int", np.dtype(np.uint32): "int"}


def _scalars(name: str, values: np.ndarray, card: int, binary: bool) -> bytes:
    """One SCALARS block; values has shape (Z, Y, X, card)."""
    t = _TYPES.get(values.dtype)
    if t is None:
        raise ValueError(f"VTK export takes float64, float32 or 32-bit integer values, not {values.dtype}")
    head = f"SCALARS {name} {t} " + (f" {card}\n" if card != 1 else "\n") + "LOOKUP_TABLE default\n"
    z, y, x = values.shape[:3]
    rows = values.reshape(z * y, x * card)
    if binary:
        body = rows.astype(values.dtype.newbyteorder(">")).tobytes() + b"\n"
    else:
        fmt = "%d " if t == "int" else "%g "
        body = "".join("".join(fmt % v for v in r.tolist()) + "\n" for r in rows).encode()
    return head.encode() + body + b"METADATA\nINFORMATION 0\n\n"


def write_vtk(file_name: str, values: np.ndarray, field_name: str, include_domain: bool = False,
              binary: bool = False, is_node_space: bool = False) -> str:
    """Write file_name + '.vtk' for values of shape (Z, Y, X) or (Z, Y, X, card)
    (x fastest); returns the path.  Mirrors Field::ioToVtk(fileName, fieldName,
    includeDomain, ioFileType, isNodeSpace) (FieldBase.h:135-140)."""
    v = np.asarray(values)
    if v.ndim == 3:
        v = v[..., None]
    if v.ndim != 4:
        raise ValueError("values must have shape (Z, Y, X) or (Z, Y, X, card)")
    z, y, x, card = v.shape
    node = (x, y, z) if is_node_space else (x + 1, y + 1, z + 1)
    vox = tuple(n - 1 for n in node)
    nvox = vox[0] * vox[1] * vox[2]
    out = [b"# vtk DataFile Version 3.0\nTitle Neon\n",
           b"BINARY\n" if binary else b"ASCII\n",
           b"DATASET STRUCTURED_POINTS\n",
           f"DIMENSIONS {node[0]} {node[1]} {node[2]}\nORIGIN 0 0 0\nSPACING 1 1 1\n".encode(),
           f"CELL_DATA {nvox}\n".encode()]
    if not is_node_space:
        out.append(_scalars(field_name, v, card, binary))
    if include_domain:
        out.append(_scalars("Domain", np.zeros((vox[2], vox[1], vox[0], 1), v.dtype), 1, binary))
    out.append(f"POINT_DATA {node[0] * node[1] * node[2]}\n".encode())
    if is_node_space:
        out.append(_scalars(field_name, v, card, binary))
    path = file_name + ".vtk"
    with open(path, "wb") as f:
        f.write(b"".join(out))
    return path


def read_vtk(path: str) -> dict:
    """Parse a file written by write_vtk / Field::ioToVtk: returns
    {"dimensions": (nx, ny, nz), "binary": bool, "cell": {name: array},
    "point": {name: array}}, arrays shaped (Z, Y, X, card)."""
    data = open(path, "rb").read()
    pos = 0

    def line() -> str:
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    if line() != "# vtk DataFile Version 3.0":
        raise ValueError("not a legacy VTK file")
    line()
    binary = line() == "BINARY"
    if line() != "DATASET STRUCTURED_POINTS":
        raise ValueError("not STRUCTURED_POINTS")
    dims = tuple(int(t) for t in line().split()[1:4])
    line()
    line()
    out = {"dimensions": dims, "binary": binary, "cell": {}, "point": {}}
    section, space = None, None
    while pos < len(data):
        s = line()
        if not s:
            continue
        w = s.split()
        if w[0] == "CELL_DATA":
            section, space = "cell", tuple(d - 1 for d in dims)
        elif w[0] == "POINT_DATA":
            section, space = "point", dims
        elif w[0] == "SCALARS":
            name, typ = w[1], w[2]
            card = int(w[3]) if len(w) > 3 else 1
            if line() != "LOOKUP_TABLE default":
                raise ValueError("missing LOOKUP_TABLE")
            dt = {"double": np.float64, "float": np.float32, "int": np.int32}[typ]
            n = space[0] * space[1] * space[2] * card
            if binary:
                nb = n * np.dtype(dt).itemsize
                arr = np.frombuffer(data[pos:pos + nb], dtype=np.dtype(dt).newbyteorder(">")).astype(dt)
                pos += nb + 1
            else:
                vals = []
                for _ in range(space[1] * space[2]):
                    vals.extend(line().split())
                arr = np.array(vals, dtype=np.float64).astype(dt)
            if arr.size != n:
                raise ValueError(f"field {name}: {arr.size} values, expected {n}")
            out[section][name] = arr.reshape(space[2], space[1], space[0], card)
            if line() != "METADATA" or line() != "INFORMATION 0":
                raise ValueError(f"field {name}: missing METADATA block")
        else:
            raise ValueError(f"unexpected line {s!r}")
    return out


def _vti_extent(box) -> str:
    return f"0 {box[0]} 0 {box[1]} 0 {box[2]}"


def _vti_array(name: str, v: np.ndarray, binary: bool) -> bytes:
    """One DataArray; v has shape (Z, Y, X, ncomp)."""
    z, y, x, nc = v.shape
    head = (f'<DataArray type="Float64" NumberOfComponents="{nc}" Name="{name}" '
            f'format="{"binary" if binary else "ascii"}">\n').encode()
    if binary:
        raw = v.astype("<f8").tobytes()
        if len(raw) > 0xFFFFFFFF:
            raise ValueError(f"binary array {name} exceeds the 4 GiB UInt32 header")
        body = base64.b64encode(len(raw).to_bytes(4, "little") + raw) + b"\n</DataArray>\n\n"
        return head + body
    fmt = "%d " if np.issubdtype(v.dtype, np.integer) else "%.17f "
    vals = v.astype(np.int64 if fmt == "%d " else np.float64).reshape(z, y, x, nc).tolist()
    parts = []
    for plane in vals:
        for row in plane:
            parts.append("".join("".join(fmt % c for c in vox) + "\t" for vox in row) + "\n")
        parts.append("\n")
    return head + "".join(parts).encode() + b"</DataArray>\n\n"


def write_vti(path: str, fields, node_space, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)) -> str:
    """Write `path` (as given) with fields = [(name, values, is_node, binary)]:
    values of shape (Z, Y, X) or (Z, Y, X, ncomp) over the node box
    node_space = (X, Y, Z) when is_node, else over the voxel box (one less in
    every direction).  Mirrors Neon::ioToVTI(gridsInfo, filename, nodeSpace,
    voxSpace, spacingData, origin) (ioToVti.h:663-719)."""
    node = tuple(int(n) for n in node_space)
    vox = tuple(n - 1 for n in node)
    arrays = {True: [], False: []}
    for name, values, is_node, binary in fields:
        v = np.asarray(values)
        if v.ndim == 3:
            v = v[..., None]
        box = node if is_node else vox
        if v.ndim != 4 or v.shape[:3] != (box[2], box[1], box[0]):
            raise ValueError(f"field {name}: shape {v.shape} does not cover the box {box} (x, y, z)")
        arrays[bool(is_node)].append(_vti_array(name, v, bool(binary)))
    f6 = lambda t: " ".join("%f" % float(a) for a in t)
    out = [b'<?xml version="1.0"?>\n',
           b'<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n',
           f'<ImageData WholeExtent="{_vti_extent(vox)} " Origin="{f6(origin)}" Spacing="{f6(spacing)} ">\n'.encode(),
           f'<Piece Extent="{_vti_extent(vox)}" >\n'.encode(),
           b"<PointData>\n", *arrays[True], b"</PointData>\n",
           b"<CellData>\n", *arrays[False], b"</CellData>\n",
           b"</Piece>\n", b" </ImageData>\n", b" </VTKFile>\n"]
    with open(path, "wb") as fh:
        fh.write(b"".join(out))
    return path


def read_vti(path: str) -> dict:
    """Parse a file written by write_vti / neon_amd::ioToVTI with an XML
    parser: returns {"whole_extent", "origin", "spacing", "point": {name:
    array}, "cell": {name: array}}, arrays shaped (Z, Y, X, ncomp) float64."""
    root = ET.parse(path).getroot()
    if root.tag != "VTKFile" or root.get("type") != "ImageData" or root.get("byte_order") != "LittleEndian":
        raise ValueError("not a little-endian ImageData VTK XML file")
    img = root.find("ImageData")
    ext = [int(t) for t in img.get("WholeExtent").split()]
    out = {"whole_extent": tuple(ext), "origin": tuple(float(t) for t in img.get("Origin").split()),
           "spacing": tuple(float(t) for t in img.get("Spacing").split()), "point": {}, "cell": {}}
    vox = (ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4])
    piece = img.find("Piece")
    for sec, tag, box in (("point", "PointData", tuple(n + 1 for n in vox)), ("cell", "CellData", vox)):
        for da in piece.find(tag).findall("DataArray"):
            if da.get("type") != "Float64":
                raise ValueError("only Float64 arrays are written")
            nc = int(da.get("NumberOfComponents"))
            n = box[0] * box[1] * box[2] * nc
            if da.get("format") == "binary":
                raw = base64.b64decode("".join(da.text.split()))
                nbytes = int.from_bytes(raw[:4], "little")
                if nbytes != 8 * n or len(raw) != 4 + nbytes:
                    raise ValueError(f"array {da.get('Name')}: header {nbytes} bytes, expected {8 * n}")
                arr = np.frombuffer(raw[4:], dtype="<f8").astype(np.float64)
            else:
                arr = np.array(da.text.split(), dtype=np.float64)
            if arr.size != n:
                raise ValueError(f"array {da.get('Name')}: {arr.size} values, expected {n}")
            out[sec][da.get("Name")] = arr.reshape(box[2], box[1], box[0], nc)
    return out


def centerlines(u: np.ndarray, ulid_x: float = 1.0):
    """Centre-line profiles of a velocity field u of shape (Z, Y, X, 3) over
    the whole box, as RunCavityTwoPop.cu:115-139 collects them: u_x along y at
    (x, z) = (X/2, Z/2) against y / Y, and u_y along x at (y, z) = (Y/2, Z/2)
    against x / X, each scaled by 1 / ulid_x.  Returns (y_profile, x_profile),
    arrays of (position, value) rows."""
    z, y, x = u.shape[:3]
    scale = 1.0 / ulid_x
    ys = np.arange(y)
    xs = np.arange(x)
    yprof = np.stack([ys / float(y), u[z // 2, :, x // 2, 0].astype(np.float64) * scale], axis=1)
    xprof = np.stack([xs / float(x), u[z // 2, y // 2, :, 1].astype(np.float64) * scale], axis=1)
    return yprof, xprof


def write_dat(path: str, rows: np.ndarray) -> None:
    """'position value' per line, printed like a default C++ stream (%g)."""
    with open(path, "w") as f:
        for a, b in rows:
            f.write("%g %g\n" % (a, b))
