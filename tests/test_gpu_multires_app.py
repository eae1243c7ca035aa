"""The C++ multi-resolution program on the Neon-shaped host API
(include/neon_amd.hpp mGrid, neon_amd/apps/lbm_multires.cpp): run with the
reference app's command line (apps/lbmMultiRes/lbmMultiRes.cu:53-95), its
centre-line files must match the reference's own output of the same run,
tests/golden/NeonMultiResLBM_5000_{X,Y}.dat (--scale 1, Re 100, KBC D3Q27,
float, 5000 iterations), and its report carries the reference's members;
the flow over a sphere (flowOverShape.h:115-190) runs and reports the same
leaves as the oracle's grid of that problem."""
import glob
import json
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "neon_amd" / "lib" / "lbm_multires"
GOLDEN = ROOT / "tests" / "golden"


def test_cli_rejects_what_is_not_built(tmp_path):
    for extra in (["--deviceType", "cpu"], ["--problemType", "mesh"], ["--dataType", "half"],
                  ["--gui"], ["--collision", "mrt"], ["--numIter", "x"]):
        r = subprocess.run([str(EXE), *extra], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage" in r.stderr, extra


@pytest.mark.gpu
def test_app_reproduces_reference_multires_output(tmp_path):
    cmd = [str(EXE), "--numIter", "5000", "--deviceType", "gpu", "--deviceId", "0", "--problemType", "lid",
           "--benchmark", "--dataType", "float", "--re", "100", "--scale", "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout)
    for name in ("NeonMultiResLBM_5000_Y.dat", "NeonMultiResLBM_5000_X.dat"):
        got_text = (tmp_path / name).read_text().split("\n")
        ref_text = (GOLDEN / name).read_text().split("\n")
        assert len(got_text) == len(ref_text)
        # the same leaves at the same positions, printed the same way
        assert [l.split(" ")[0] for l in got_text] == [l.split(" ")[0] for l in ref_text]
        got, ref = np.loadtxt(tmp_path / name), np.loadtxt(GOLDEN / name)
        dev = np.abs(got[:, 1] - ref[:, 1])
        print(f"{name}: max |u - u_ref| / U {dev.max():.3e}, median {np.median(dev):.3e}")
        assert dev.max() < 0.005
    reports = glob.glob(str(tmp_path / "MultiResLBM_P1_CH-SEO+_F_D*.json"))
    assert len(reports) == 1
    rep = json.load(open(reports[0]))
    assert rep["Depth"] == 3 and rep["Collision"] == "KBC" and rep["NumIter"] == 5000
    assert rep["Grid Size X"] == 160 and rep["clength"] == 40 and rep["Algorithm"] == "CH-SEO+"
    assert len(rep["NumActiveVoxels"]) == 3 and rep["MLUPS"] > 0 and rep["EMLUPS"] > rep["MLUPS"]
    assert abs(rep["omega"] - 1.0 / (3 * 0.04 * 40 / 100 + 0.5)) < 1e-6


@pytest.mark.gpu
def test_app_runs_the_flow_over_a_sphere(tmp_path):
    import neon_amd.mgrid as mg
    from oracle.oracle import MresOracle
    cmd = [str(EXE), "--numIter", "40", "--deviceType", "gpu", "--deviceId", "0", "--problemType", "sphere",
           "--benchmark", "--dataType", "float", "--re", "100", "--scale", "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout)
    assert not glob.glob(str(tmp_path / "*.dat"))  # no centre lines for the sphere
    reports = glob.glob(str(tmp_path / "MultiResLBM_P1_CH-SEO+_F_D*.json"))
    assert len(reports) == 1
    rep = json.load(open(reports[0]))
    assert rep["Record Name"] == "Sphere MultiRes LBM" and rep["problemType"] == "sphere"
    assert (rep["Grid Size X"], rep["Grid Size Y"], rep["Grid Size Z"]) == (136, 96, 136)
    assert rep["Depth"] == 3 and rep["clength"] == 2 and rep["Collision"] == "KBC"
    uin, om0 = mg.sphere_parameters(1, 100.0, "f32")
    assert abs(rep["omega"] - om0) < 1e-6
    dim, lams, _, _ = mg.sphere_problem(1)
    o = MresOracle(dim, lams)
    leaves = [int((o.active(l) & ~o.refined(l)).sum()) if l else int(o.active(0).sum()) for l in range(3)]
    assert rep["NumActiveVoxels"] == leaves


def _read_vtk(path):
    """(points, cells, cell types, velocity, level) of the unstructured-grid
    file postProcess writes."""
    lines = path.read_text().split("\n")
    assert lines[:4] == ["# vtk DataFile Version 2.0", "mGrid", "ASCII", "DATASET UNSTRUCTURED_GRID"]
    k = 4
    npts = int(lines[k].split()[1]); k += 1
    pts = np.array([[float(v) for v in lines[k + i].split()] for i in range(npts)]); k += npts
    ncell = int(lines[k].split()[1]); assert int(lines[k].split()[2]) == 9 * ncell; k += 1
    cells = np.array([[int(v) for v in lines[k + i].split()] for i in range(ncell)]); k += ncell
    assert lines[k].split()[:2] == ["CELL_TYPES", str(ncell)]; k += 1
    types = np.array([int(lines[k + i]) for i in range(ncell)]); k += ncell
    assert lines[k].split() == ["CELL_DATA", str(ncell)]; k += 1
    assert lines[k].split() == ["VECTORS", "Velocity", "float"]; k += 1
    vel = np.array([[float(v) for v in lines[k + i].split()] for i in range(ncell)]); k += ncell
    assert lines[k].split() == ["SCALARS", "Level", "float", "1"] and lines[k + 1].split() == ["LOOKUP_TABLE", "default"]
    k += 2
    level = np.array([int(lines[k + i]) for i in range(ncell)])
    return pts, cells, types, vel, level


@pytest.mark.gpu
def test_app_vtk_output(tmp_path):
    """--visual --vtk --freq 50 --numIter 100 on the --scale 1 lid grid: the
    files of iterations 50 and 100 (postProcess.h:94-151) hold every leaf on
    the z mid-plane (the default --sliceZ 1) as a VTK voxel of side 2^level,
    and the velocities of the leaves on the centre line x = z = 80 equal the
    centre-line file of the same run."""
    cmd = [str(EXE), "--numIter", "100", "--deviceType", "gpu", "--problemType", "lid", "--visual", "--vtk",
           "--freq", "50", "--dataType", "float", "--re", "100", "--scale", "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(p.name for p in tmp_path.glob("Velocity_*.vtk")) == ["Velocity_0050.vtk", "Velocity_0100.vtk"]
    pts, cells, types, vel, level = _read_vtk(tmp_path / "Velocity_0100.vtk")
    assert (types == 11).all() and cells[:, 0].tolist() == [8] * len(cells)
    assert set(level.tolist()) == {0, 1, 2}
    corners = pts[cells[:, 1:]]                      # n x 8 x 3
    lo, side = corners[:, 0, :], (2.0 ** level)[:, None]
    # the reference's corner order (postProcess.h:111-126 over :283-321): x
    # fastest, then z, then y
    for j, (dx, dy, dz) in enumerate([(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1),
                                      (0, 1, 0), (1, 1, 0), (0, 1, 1), (1, 1, 1)]):
        assert np.array_equal(corners[:, j, :], lo + side * np.array([dx, dy, dz]))
    # every drawn leaf reaches the mid-plane z = 80
    assert ((lo[:, 2] <= 80) & (lo[:, 2] + side[:, 0] >= 80)).all()
    # the centre line x = z = 80: the same velocities as the .dat file
    dat = np.loadtxt(tmp_path / "NeonMultiResLBM_0100_Y.dat")
    on = (lo[:, 0] == 80) & (lo[:, 2] == 80)
    got = sorted(zip(lo[on, 1] / 160.0, vel[on, 0] / np.float32(0.04)))
    assert len(got) == len(dat)
    assert np.allclose(np.array(got), dat, rtol=1e-5, atol=1e-7)
    # no slice: every leaf of every level, as many as the report counts
    whole = tmp_path / "whole"
    whole.mkdir()
    r = subprocess.run([str(EXE), "--numIter", "1", "--deviceType", "gpu", "--problemType", "lid", "--visual",
                        "--vtk", "--sliceZ", "-1", "--dataType", "float", "--re", "100", "--scale", "1"],
                       cwd=whole, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    _, cells1, _, vel1, level1 = _read_vtk(whole / "Velocity_0001.vtk")
    rep = json.load(open(glob.glob(str(whole / "MultiResLBM_P1_*.json"))[0]))
    assert len(cells1) == sum(rep["NumActiveVoxels"])
    assert np.bincount(level1).tolist() == rep["NumActiveVoxels"]


@pytest.mark.gpu
def test_app_binary_output(tmp_path):
    """--visual --binary (postProcess.h:153-227): the velocity sampled on the
    top level's lattice (40^3 for the --scale 1 grid, every leaf whose origin
    is a multiple of 4) in a binary legacy VTK file; along the centre line
    x = z = 80 it equals the centre-line file of the same run."""
    cmd = [str(EXE), "--numIter", "100", "--deviceType", "gpu", "--problemType", "lid", "--visual", "--binary",
           "--freq", "50", "--dataType", "float", "--re", "100", "--scale", "1"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(p.name for p in tmp_path.glob("BinVelocity_*.vtk")) == ["BinVelocity_0050.vtk",
                                                                          "BinVelocity_0100.vtk"]
    raw = (tmp_path / "BinVelocity_0100.vtk").read_bytes()
    head = raw[:raw.index(b"LOOKUP_TABLE default\n")].decode().split("\n")
    assert head[:4] == ["# vtk DataFile Version 3.0", "Title Neon", "BINARY", "DATASET STRUCTURED_POINTS"]
    assert head[4] == "DIMENSIONS 41 41 41" and head[7] == "CELL_DATA 64000"
    assert head[8].split() == ["SCALARS", "V", "float", "3"]
    start = raw.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    v = np.frombuffer(raw[start:start + 64000 * 12], dtype=">f4").reshape(40, 40, 40, 3)   # z, y, x, c
    assert np.isfinite(v).all()
    dat = np.loadtxt(tmp_path / "NeonMultiResLBM_0100_Y.dat")
    checked = 0
    for pos, val in dat:
        y = int(round(pos * 160))
        if y % 4 == 0:
            assert abs(v[20, y // 4, 20, 0] / np.float32(0.04) - val) <= 1e-5 * abs(val) + 1e-7, (y, val)
            checked += 1
    assert checked >= 30
