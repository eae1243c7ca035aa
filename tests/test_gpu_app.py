"""The C++ cavity program on the Neon-shaped host API (include/neon_amd.hpp,
neon_amd/apps/lbm_cavity.cpp): run with the reference benchmark's command
line, its centre-line files must match the oracle run of the same case."""
import json
import pathlib
import subprocess

import numpy as np
import pytest

EXE = pathlib.Path(__file__).resolve().parents[1] / "neon_amd" / "lib" / "lbm_cavity"


def run_app(tmp, *args):
    cmd = [str(EXE), "--deviceType", "gpu", "--deviceIds", "0", *map(str, args)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cli_rejects_what_is_not_built(tmp_path):
    for extra in (["--grid", "eGrid"], ["--deviceType", "cpu"], ["--storageFP", "quad"],
                  ["--stagedNeighbourhood"], ["--grid", "bGrid", "--stagedNeighbourhood", "--storageFP", "double"]):
        r = subprocess.run([str(EXE), "--deviceType", "gpu", "--deviceIds", "0", *extra],
                           cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([str(EXE), "--deviceType", "gpu", "--deviceIds", "0", "1"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "one device id" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("grid,storage,compute", [("dGrid", "double", "double"),
                                                  ("bGrid", "double", "double"),
                                                  ("dGrid", "float", "float"),
                                                  ("bGrid", "float", "float")])
def test_app_centerline_matches_oracle(orc, tmp_path, grid, storage, compute):
    N, Re, ulb = 24, 100.0, 0.04
    run_app(tmp_path, "--grid", grid, "--domain-size", N, "--storageFP", storage, "--computeFP", compute,
            "--visual", "--max-iter", 101, "--Re", Re, "--ulb", ulb)
    got = np.loadtxt(tmp_path / "NeonUniformLBM_00100_Y.dat")
    nu = ulb * (N - 2) / Re
    omega = 1.0 / (3.0 * nu + 0.5)
    st, cm = (orc.F64, orc.COMPUTE_F64) if storage == "double" else (orc.F32, orc.COMPUTE_REF)
    o = orc.DenseOracle((N, N, N), st, cm)
    o.setup_cavity(ulb)
    o.run(100, omega)
    _, u = o.moments()
    ref = u[0, N // 2, :, N // 2].astype(np.float64)        # u_x along y at x = z = N/2
    assert np.allclose(got[:, 0], np.arange(N) / N)
    # six significant digits in the file
    assert np.allclose(got[:, 1], ref, rtol=2e-5, atol=1e-9), np.abs(got[:, 1] - ref).max()
    gx = np.loadtxt(tmp_path / "NeonUniformLBM_00100_X.dat")
    assert np.allclose(gx[:, 1], u[1, N // 2, N // 2, :], rtol=2e-5, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("grid,storage,compute", [("dGrid", "float", "float-fast"),
                                                  ("bGrid", "half-dev", "float"),
                                                  ("dGrid", "float-dev", "float")])
def test_app_benchmark_mode_reports(tmp_path, grid, storage, compute):
    out = run_app(tmp_path, "--grid", grid, "--domain-size", 64, "--storageFP", storage,
                  "--computeFP", compute, "--benchmark", "--warmup-iter", 10, "--max-iter", 60,
                  "--report-filename", "rep")
    assert "MLUPS" in out and "Warm up completed (10 iterations )" in out
    rep = json.loads((tmp_path / "rep.json").read_text())
    assert rep["gridType"] == grid and rep["N"] == 64 and rep["MLUPS"][0] > 0
    assert abs(rep["omega"] - 1.0 / (3.0 * 0.04 * 62 / 100.0 + 0.5)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("grid", ["dGrid", "bGrid"])
def test_app_vtk_export_matches_oracle(orc, tmp_path, grid):
    from neon_amd import export
    N, Re, ulb = 12, 100.0, 0.04
    run_app(tmp_path, "--grid", grid, "--domain-size", N, "--storageFP", "double", "--computeFP", "double",
            "--visual", "--vti", "--max-iter", 101, "--Re", Re, "--ulb", ulb)
    for it in ("00000", "00100"):
        assert (tmp_path / f"u_{it}.vtk").exists() and (tmp_path / f"rho_{it}.vtk").exists()
    u_file = export.read_vtk(str(tmp_path / "u_00100.vtk"))
    rho_file = export.read_vtk(str(tmp_path / "rho_00100.vtk"))
    assert u_file["dimensions"] == (N + 1, N + 1, N + 1) and not u_file["binary"]
    u = u_file["cell"]["u"]
    rho = rho_file["cell"]["rho"][..., 0]
    assert u.shape == (N, N, N, 3) and u.dtype == np.float64
    nu = ulb * (N - 2) / Re
    o = orc.DenseOracle((N, N, N), orc.F64, orc.COMPUTE_F64)
    o.setup_cavity(ulb)
    o.run(100, 1.0 / (3.0 * nu + 0.5))
    r_ref, u_ref = o.moments()
    b = (slice(1, N - 1),) * 3                      # bulk cells
    # six significant digits in the text file
    np.testing.assert_allclose(rho[b], r_ref[b], rtol=6e-6, atol=0)
    for k in range(3):
        np.testing.assert_allclose(u[b + (k,)], u_ref[(k,) + b], rtol=2e-5, atol=1e-10)
    # the centre line written beside it is the same field
    y = np.loadtxt(tmp_path / "NeonUniformLBM_00100_Y.dat")
    np.testing.assert_allclose(y[:, 1], u[N // 2, :, N // 2, 0], rtol=2e-5, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("storage,compute", [("float", "float"), ("half-dev", "float")])
def test_app_staged_neighbourhood_writes_the_same_files(tmp_path, storage, compute):
    """lbm_cavity --grid bGrid with and without --stagedNeighbourhood (the
    LDS-staged step through bGrid::setStagedNeighbourhood): the centre-line
    files are identical byte for byte."""
    out = {}
    for staged in (False, True):
        d = tmp_path / ("staged" if staged else "plain")
        d.mkdir()
        extra = ["--stagedNeighbourhood"] if staged else []
        run_app(d, "--grid", "bGrid", "--domain-size", 40, "--storageFP", storage, "--computeFP", compute,
                "--visual", "--max-iter", 101, "--Re", 100.0, "--ulb", 0.04, *extra)
        out[staged] = [(d / f"NeonUniformLBM_00100_{a}.dat").read_bytes() for a in ("X", "Y")]
    assert out[False] == out[True]
