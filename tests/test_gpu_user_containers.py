"""A caller's own containers on the Neon-shaped API (grid.newContainer,
Loader::load, partition getNghData; neon_amd/apps/user_containers.hip,
compiled with hipcc against include/neon_amd.hpp): the D3Q19 cavity step
written as user code agrees bit for bit with the library's
LbmIterationD3Q19 (binary64), a 7-point STENCIL average agrees with the
host, and a STENCIL load of a field the container writes is refused."""
import pathlib
import subprocess

import pytest

EXE = pathlib.Path(__file__).resolve().parents[1] / "neon_amd" / "lib" / "user_containers"


def test_usage_without_device_work():
    r = subprocess.run([str(EXE), "40", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [(40, 30), (23, 12)])
def test_user_containers(n, steps):
    print(f"user_containers {n} {steps}", flush=True)
    r = subprocess.run([str(EXE), str(n), str(steps)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{n}^3 cavity, {steps} steps: 0 of " in r.stdout
    assert "user containers OK" in r.stdout
