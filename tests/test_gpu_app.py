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
    for extra in (["--grid", "eGrid"], ["--vti"], ["--storageFP", "quad"]):
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
