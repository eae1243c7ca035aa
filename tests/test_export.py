"""Exports in the reference's on-disk formats (CPU).

* Legacy VTK: the C++ writer behind Field::ioToVtk (include/neon_amd.hpp) and
  neon_amd.export.write_vtk must write identical bytes; the layout is pinned
  by a hand-written file for a 2x1x1 field that follows
  libNeonCore/include/Neon/core/tools/io/ioToVTK.h:194-272, :320-354 and
  :413-445 line by line; read_vtk round-trips both encodings.
* Centre-line .dat files: the reference's own output
  (apps/lbmMultiRes/scripts/NeonUniformLBM_20000_Y.dat, copied as
  tests/golden/NeonUniformLBM_20000_Y.dat) parses and re-prints unchanged
  through write_dat, which prints like a default C++ stream.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from neon_amd import export

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("vtk") / "vtk_writer")
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "vtk_writer.cpp"), "-o", exe], check=True)
    return exe


def _values(nx, ny, nz, card, dtype):
    z, y, x, k = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(card), indexing="ij")
    v = (x + 10 * y + 100 * z).astype(np.float64) / 7.0 + k.astype(np.float64) / 1000.0
    return v.astype(dtype)


CASES = [
    (3, 2, 2, 1, "double", 0, 0, 0),
    (3, 2, 2, 3, "double", 0, 1, 0),
    (4, 3, 2, 3, "float", 0, 0, 0),
    (4, 3, 2, 1, "int", 0, 0, 1),
    (3, 2, 2, 3, "double", 1, 0, 0),
    (4, 3, 2, 1, "float", 1, 1, 0),
    (2, 2, 3, 3, "float", 1, 0, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[4]}-c{c[3]}-b{c[5]}-d{c[6]}-n{c[7]}" for c in CASES])
def test_cpp_and_python_writers_agree(writer, tmp_path, case):
    nx, ny, nz, card, typ, binary, domain, node = case
    dt = {"double": np.float64, "float": np.float32, "int": np.int32}[typ]
    cpp = str(tmp_path / "cpp")
    subprocess.run([writer, cpp, str(nx), str(ny), str(nz), str(card), typ, str(binary), str(domain), str(node)],
                   check=True)
    v = _values(nx, ny, nz, card, dt)
    py = export.write_vtk(str(tmp_path / "py"), v, "field", include_domain=bool(domain), binary=bool(binary),
                          is_node_space=bool(node))
    a = open(cpp + ".vtk", "rb").read()
    b = open(py, "rb").read()
    assert a == b
    back = export.read_vtk(py)
    sec = "point" if node else "cell"
    got = back[sec]["field"]
    assert got.dtype == dt and got.shape == (nz, ny, nx, card)
    if binary or typ == "int":
        np.testing.assert_array_equal(got, v)
    else:
        np.testing.assert_allclose(got, v, rtol=6e-6, atol=0)  # six significant digits
    if domain:
        assert not back["cell"]["Domain"].any()
    node_dims = (nx, ny, nz) if node else (nx + 1, ny + 1, nz + 1)
    assert back["dimensions"] == node_dims


def test_layout_matches_reference_writer(tmp_path):
    """A 2x1x1 two-component double field, ASCII, cell data (isNodeSpace
    false, the default of Field::ioToVtk), written out by hand from the
    reference's writer."""
    v = np.array([[[[0.5, -1.25], [3.0, 1e-7]]]])
    path = export.write_vtk(str(tmp_path / "u_00000"), v, "u")
    want = ("# vtk DataFile Version 3.0\n"
            "Title Neon\n"
            "ASCII\n"
            "DATASET STRUCTURED_POINTS\n"
            "DIMENSIONS 3 2 2\n"
            "ORIGIN 0 0 0\n"
            "SPACING 1 1 1\n"
            "CELL_DATA 2\n"
            "SCALARS u double  2\n"
            "LOOKUP_TABLE default\n"
            "0.5 -1.25 3 1e-07 \n"
            "METADATA\n"
            "INFORMATION 0\n"
            "\n"
            "POINT_DATA 12\n")
    assert open(path).read() == want


def test_binary_is_big_endian(tmp_path):
    v = np.array([[[1.0]]], dtype=np.float32)
    data = open(export.write_vtk(str(tmp_path / "b"), v, "rho", binary=True), "rb").read()
    body = data.split(b"LOOKUP_TABLE default\n", 1)[1]
    assert body[:4] == b"\x3f\x80\x00\x00"  # 1.0f, most significant byte first
    assert body[4:5] == b"\n"


def test_rejects_unsupported_types(tmp_path):
    with pytest.raises(ValueError):
        export.write_vtk(str(tmp_path / "h"), np.zeros((1, 1, 1), np.float16), "h")


def test_golden_centerline_reprints_unchanged(tmp_path):
    src = os.path.join(HERE, "golden", "NeonUniformLBM_20000_Y.dat")
    rows = np.loadtxt(src)
    out = str(tmp_path / "Y.dat")
    export.write_dat(out, rows)
    assert open(out).read() == open(src).read()


def test_centerlines_positions_and_components():
    n = 8
    u = np.zeros((n, n, n, 3))
    u[n // 2, :, n // 2, 0] = np.arange(n) * 0.01
    u[n // 2, n // 2, :, 1] = -np.arange(n) * 0.02
    y, x = export.centerlines(u, ulid_x=1.0)
    np.testing.assert_array_equal(y[:, 0], np.arange(n) / n)
    np.testing.assert_array_equal(y[:, 1], np.arange(n) * 0.01)
    np.testing.assert_array_equal(x[:, 1], -np.arange(n) * 0.02)


# ---- XML ImageData (.vti) ---------------------------------------------------

@pytest.fixture(scope="module")
def vti_writer(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("vti") / "vti_writer")
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "vti_writer.cpp"), "-o", exe], check=True)
    return exe


def _vti_values(box, nc, dtype):
    x, y, z = box
    zz, yy, xx, k = np.meshgrid(np.arange(z), np.arange(y), np.arange(x), np.arange(nc), indexing="ij")
    v = (xx + 10 * yy + 100 * zz).astype(np.float64) / 7.0 + k.astype(np.float64) / 1000.0 - 3.0
    return v.astype(dtype)


VTI_CASES = [((3, 2, 2), "double", 0), ((4, 3, 5), "double", 7), ((4, 3, 2), "float", 5),
             ((5, 2, 3), "int", 2), ((2, 2, 2), "double", 1)]


@pytest.mark.parametrize("node,typ,mask", VTI_CASES, ids=[f"{c[1]}-m{c[2]}-{c[0][0]}{c[0][1]}{c[0][2]}" for c in VTI_CASES])
def test_vti_cpp_and_python_writers_agree(vti_writer, tmp_path, node, typ, mask):
    dt = {"double": np.float64, "float": np.float32, "int": np.int32}[typ]
    cpp = str(tmp_path / "cpp.vti")
    subprocess.run([vti_writer, cpp, *map(str, node), typ, str(mask)], check=True)
    vox = tuple(n - 1 for n in node)
    u, rho, w = _vti_values(vox, 3, dt), _vti_values(node, 1, dt), _vti_values(vox, 2, dt)
    py = export.write_vti(str(tmp_path / "py.vti"),
                          [("u", u, False, mask & 1), ("rho", rho, True, mask & 2), ("w", w, False, mask & 4)],
                          node, spacing=(0.5, 1.0, 2.0), origin=(-1.0, 0.0, 3.25))
    assert open(cpp, "rb").read() == open(py, "rb").read()
    back = export.read_vti(py)
    assert back["whole_extent"] == (0, vox[0], 0, vox[1], 0, vox[2])
    assert back["origin"] == (-1.0, 0.0, 3.25) and back["spacing"] == (0.5, 1.0, 2.0)
    for name, want, sec, binary in (("u", u, "cell", mask & 1), ("rho", rho, "point", mask & 2),
                                    ("w", w, "cell", mask & 4)):
        got = back[sec][name]
        if binary or typ != "double":
            np.testing.assert_array_equal(got, want.astype(np.float64))
        else:
            np.testing.assert_allclose(got, want, rtol=1e-15, atol=1e-16)  # 17 decimals


def test_vti_layout_matches_reference_writer(tmp_path):
    """Two voxels along x (node box 3x2x2), a two-component cell field in
    ASCII, written out by hand from the reference's writer
    (ioToVti.h:431-462 for the array, :695-718 for the file around it)."""
    v = np.array([[[[0.5, -1.25], [3.0, 1e-7]]]])
    path = export.write_vti(str(tmp_path / "a.vti"), [("u", v, False, False)], (3, 2, 2))
    want = ('<?xml version="1.0"?>\n'
            '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n'
            '<ImageData WholeExtent="0 2 0 1 0 1 " Origin="0.000000 0.000000 0.000000" '
            'Spacing="1.000000 1.000000 1.000000 ">\n'
            '<Piece Extent="0 2 0 1 0 1" >\n'
            '<PointData>\n'
            '</PointData>\n'
            '<CellData>\n'
            '<DataArray type="Float64" NumberOfComponents="2" Name="u" format="ascii">\n'
            '0.50000000000000000 -1.25000000000000000 \t3.00000000000000000 0.00000010000000000 \t\n'
            '\n'
            '</DataArray>\n\n'
            '</CellData>\n'
            '</Piece>\n'
            ' </ImageData>\n'
            ' </VTKFile>\n')
    assert open(path).read() == want


def test_vti_binary_header_and_encoding(tmp_path):
    """One node value 1.0 in binary: base64 of the UInt32 byte count 8 and the
    little-endian binary64 bytes of 1.0, as one stream."""
    path = export.write_vti(str(tmp_path / "b.vti"), [("r", np.ones((1, 1, 1)), True, True)], (1, 1, 1))
    body = open(path).read().split('format="binary">\n', 1)[1].split("\n", 1)[0]
    import base64
    assert base64.b64decode(body) == b"\x08\x00\x00\x00" + b"\x00" * 6 + b"\xf0\x3f"


def test_vti_rejects_bad_boxes(vti_writer, tmp_path):
    with pytest.raises(ValueError):  # cell values over the node box
        export.write_vti(str(tmp_path / "c.vti"), [("u", np.zeros((2, 2, 2)), False, False)], (2, 2, 2))
    r = subprocess.run([vti_writer, str(tmp_path / "d.vti"), "0", "2", "2", "double", "1"], capture_output=True)
    assert r.returncode == 1 and b"negative voxel box" in r.stderr
