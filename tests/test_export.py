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
