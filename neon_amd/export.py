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
* Centre lines (``centerlines`` / ``write_dat``): the two
  NeonUniformLBM_<iter>_{X,Y}.dat files of the cavity benchmark
  (benchmarks/lbm-lid-driven-cavity-flow/src/RunCavityTwoPop.cu:113-149).
"""
from __future__ import annotations

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
