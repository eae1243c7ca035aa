"""Run the CPU oracle on the reference's own cavity case and write its
centre-line profile, for comparison with the reference-held output
NeonUniformLBM_20000_Y.dat.

Case (benchmarks/lbm-lid-driven-cavity-flow/src/app.cpp:15-17, Config.h:19-41,
Config.cpp:105-111): N = 160, Re = 100, ulb = 0.04, nu = ulb (N-2) / Re,
omega = 1 / (3 nu + 1/2), dGrid, double storage and compute, 20000 steps.
Export (RunCavityTwoPop.cu:82-149): computeRhoAndU on the current field,
u_x at x = N/2, z = N/2 for every y, written as "y/N u_x".

    python tests/golden/make_centerline.py [--threads T]

Writes tests/golden/oracle_centerline_160_20000.txt (about 50 min on 8 cores).
"""
import argparse
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle as orc  # noqa: E402

N, RE, ULB, STEPS = 160, 100.0, 0.04, 20000


def omega_of(n=N, re=RE, ulb=ULB):
    nu = ulb * (n - 2) / re
    return 1.0 / (3.0 * nu + 0.5)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        orc.set_threads(a.threads)
    o = orc.DenseOracle((N, N, N))
    o.setup_cavity(ULB)
    t0 = time.time()
    chunk = 500
    for s in range(0, STEPS, chunk):
        o.run(chunk, omega_of())
        print(f"step {s + chunk} / {STEPS}  {time.time() - t0:.0f} s", flush=True)
    _, u = o.moments()
    line = u[0][N // 2, :, N // 2]          # u[a][z, y, x]
    out = HERE / "oracle_centerline_160_20000.txt"
    with open(out, "w") as f:
        f.write(f"# oracle: N={N} Re={RE} ulb={ULB} steps={STEPS} omega={omega_of()!r}\n")
        for y in range(N):
            f.write(f"{y / N!r} {float(line[y])!r}\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
