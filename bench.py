"""Benchmark of the D3Q19 lattice-Boltzmann step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload: lid-driven cavity, D3Q19 BGK, dense 256^3 grid (254^3 fluid
cells), FP32 storage and FP32 collision, Re = 1000 (nu = U N / Re, U = 0.05),
W = 100 warm-up steps then K = 500 timed steps.  A step is one launch of the
fused pull-stream + bounce-back + BGK kernel over the whole grid; inputs are
resident in HBM when the timed region starts.

Prints ONE JSON line: MLUPS over fluid cells (value), the roofline of the step
kernel (algorithmic bytes 2*19*sizeof(storage) per fluid cell / average launch
time from HIP events on the launch stream, vs 8 TB/s), the rate of the same
step with the reference's float-path collision (mlups_compute_ref), the
velocity error against a binary64 GPU run of the same case at a fixed horizon
of 600 steps from rest (whatever --steps is), and two CPU baselines on a
bounded sample of the same workload: cpu_baseline, the CPU oracle on one
thread in the reference's serial z-y-x order (Neon's CPU launcher has its
OpenMP pragmas commented out), and cpu_baseline_all_cores, the same oracle
with OpenMP on every host thread.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "D3Q19 MLUPS & %HBM roofline; time-to-steady warm-start vs plain LBM"
HBM_PEAK_GBS = 8000.0
STORAGE_BYTES = {"f64": 8, "f32": 4, "f32dev": 4, "f16dev": 2}

WORKLOADS = {
    # name: (N, storage, compute, Re, default steps, default warmup)
    "cavity256_f32": (256, "f32", "f32", 1000.0, 500, 100),
    "cavity256_f16dev": (256, "f16dev", "f32", 1000.0, 500, 100),
    "cavity256_f32dev": (256, "f32dev", "f32", 1000.0, 500, 100),
    "cavity256_f64": (256, "f64", "f64", 1000.0, 200, 20),
    "cavity64_f64": (64, "f64", "f64", 100.0, 1000, 0),
}

# Porous medium on the block-sparse grid: 384^3, 8^3 blocks, solid = union of
# 2000 seeded spheres with radius U[6,20), fully solid blocks inactive, body
# force g = 1e-6 along x, periodic x, nu = 0.02.
POROUS_ERROR_STEPS = 220   # horizon of the porous velocity error (the default 20 warm-up + 200 timed)
POROUS = {
    # name: (N, storage, compute, default steps, default warmup)
    "porous384_f32": (384, "f32", "f32", 200, 20),
    "porous384_f16dev": (384, "f16dev", "f32", 200, 20),
    # the same step with the block neighbourhood staged in LDS (one workgroup
    # per 2 x 2 x 1 tile of blocks, nx_d3q19_step_block_tiled)
    "porous384_f32_tiled": (384, "f32", "f32", 200, 20, "tiled"),
    "porous384_f16dev_tiled": (384, "f16dev", "f32", 200, 20, "tiled"),
    # 2-byte storage, two x-adjacent voxels per lane (nx_d3q19_set_block_pairs)
    "porous384_f16dev_pairs": (384, "f16dev", "f32", 200, 20, "pairs"),
}


# Batched cavities: 512 independent lid-driven cubes, edge U{32..96} step 8,
# Re U[50,1000], lid 0.05, 0-4 seeded boxes with edges U{4..16}; one
# persistent launch steps every member.
BATCH = {
    # name: (members, storage, default steps, default warmup)
    "batch512_f32": (512, "f32", 200, 20),
    "batch512_f16dev": (512, "f16dev", 200, 20),
}


# The same 512 batched cavities run to steady state: FP32 deviation storage
# (binary32 f - w_q; raw binary32 populations stall near a residual of 1e-5 on
# small cavities), nx_batch_run to r = ||u_t - u_{t-K}|| / ||u_t|| < 1e-6 with
# K = 100 (FP64 sums over bulk cells), converged members retire; capped at a
# stated number of steps per member (--steps overrides the cap).
CONVERGE = {
    # name: (members, storage, tol, check interval K, default step cap)
    "batch512_converge": (512, "f32dev", 1e-6, 100, 100000),
}


# Refined lid-driven cavity (the lbmMultiRes lid problem, BGK D3Q19).
# mres256_*: 256^3 domain, two levels, level 0 = the shell within a quarter
# of the half width of the faces (sdfCube > -1/4), level 1 the interior, Re
# 1000.  mres_scale*: the app's own --scale configurations
# (apps/lbmMultiRes/lidDrivenCavity.h:95-214; grid edge and levelSDF, depth
# 3 or 4) at Re 100, as its run.sh sweeps them.  ulb 0.04, Re on the top
# level's length.  A step is one coarse step: the level recursion, level l
# updated 2^(depth-1-l) times.
MRES = {
    # name: (N, levelSDF, storage, default steps, default warmup, Re[, lattice]);
    # the sphere entries: ("sphere", --scale, storage, steps, warmup, Re[, lattice])
    "mres256_f32": (256, (0.0, -0.25, -1.0), "f32", 100, 10, 1000.0),
    "mres256_f64": (256, (0.0, -0.25, -1.0), "f64", 50, 5, 1000.0),
    "mres_scale2_f32": (240, (0.0, -24.0 / 120.0, -80.0 / 120.0, -1.0), "f32", 50, 5, 100.0),
    "mres_scale2_f64": (240, (0.0, -24.0 / 120.0, -80.0 / 120.0, -1.0), "f64", 25, 3, 100.0),
    "mres_scale5_f32": (512, (0.0, -64.0 / 256.0, -112.0 / 256.0, -1.0), "f32", 20, 2, 100.0),
    "mres_scale9_f32": (480, (0.0, -48.0 / 240.0, -96.0 / 240.0, -160.0 / 240.0, -1.0), "f32", 10, 1, 100.0),
    # the app's default build: KBC collision on D3Q27
    "mres_scale2_kbc_f32": (240, (0.0, -24.0 / 120.0, -80.0 / 120.0, -1.0), "f32", 50, 5, 100.0, "kbc27"),
    "mres_scale2_kbc_f64": (240, (0.0, -24.0 / 120.0, -80.0 / 120.0, -1.0), "f64", 25, 3, 100.0, "kbc27"),
    "mres_scale9_kbc_f32": (480, (0.0, -48.0 / 240.0, -96.0 / 240.0, -160.0 / 240.0, -1.0), "f32", 10, 1, 100.0,
                            "kbc27"),
    # the app's flow over a sphere (flowOverShape.h:115-190) at --scale 2:
    # 272 x 192 x 272, three levels, inlet u = 0.04 at x = 0, Re 100 on the
    # sphere's radius / 4, the app's default KBC build in float
    "mres_sphere2_kbc_f32": ("sphere", 2, "f32", 50, 5, 100.0, "kbc27"),
    "mres_sphere4_kbc_f32": ("sphere", 4, "f32", 20, 2, 100.0, "kbc27"),
    "mres_sphere2_f32": ("sphere", 2, "f32", 50, 5, 100.0),
}


# Duct with Zou-He ends: X x Y x Y cells along x, bounce-back walls on the
# four side faces, density faces rho_in / rho_out on the x planes (the
# flow-over-sphere benchmark's inlet/outlet cells), nu = 0.02; stepped by
# nx_d3q19_step_dense_zouhe (one thread per cell, flag word + pull).
ZOUHE = {
    # name: (X, Y, storage, default steps, default warmup)
    "duct512_zouhe_f32": (512, 128, "f32", 200, 20),
    "duct512_zouhe_f64": (512, 128, "f64", 100, 10),
}


def dense_kernel_name(t: dict) -> str:
    """Container step kernel that a dense launch configuration runs."""
    if t.get("vx", 1) > 1:
        return "k_fused_vx"
    return {0: "k_fused", 1: "k_fused_fix", 2: "k_fused_fix[bitmap]", 3: "k_fused_edge"}[t.get("boundary", 0)]


def traffic_from_profiles(kernel_tag: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel of known size), if committed."""
    p = ROOT / "profiles" / f"traffic_{kernel_tag}.json"
    if not p.exists():
        return None
    try:
        return float(json.loads(p.read_text())["bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(N, storage, compute, omega, budget_s=12.0, max_steps=50, threads=0):
    """Oracle (the C restatement of the reference step) on the host, a few
    steps of the same grid.  threads=1: one host thread walking z, y, x in
    order -- the reference's CPU launcher, whose OpenMP pragmas are commented
    out (libNeonSet/include/Neon/set/LambdaExecutor.h:80-96, :119-144);
    threads=0: every host thread (OpenMP over z slabs)."""
    from oracle import oracle as orc
    st = {"f64": orc.F64, "f32": orc.F32, "f32dev": orc.F32DEV, "f16dev": orc.F16DEV}[storage]
    cm = {"f64": orc.COMPUTE_F64, "f32": orc.COMPUTE_F32, "ref": orc.COMPUTE_REF}[compute]
    orc.set_threads(threads)
    try:
        o = orc.DenseOracle((N, N, N), st, cm)
        o.setup_cavity(0.05)
        o.run(1, omega)                       # first touch
        steps, t0 = 0, time.perf_counter()
        while True:
            o.run(1, omega)
            steps += 1
            el = time.perf_counter() - t0
            if el > budget_s or steps >= max_steps:
                break
        cores = int(orc.lib.or_get_threads())
    finally:
        orc.set_threads(0)
    fluid = (N - 2) ** 3
    how = ("one thread, z-y-x loop as the reference's serial CPU launcher" if threads == 1
           else "OpenMP over z, all host threads")
    return {"value": fluid * steps / el / 1e6, "unit": "MLUPS", "cores": cores,
            "kind": "port",
            "sample": f"{steps} steps of the {N}^3 cavity ({storage} storage, {compute} compute), "
                      f"{el:.1f} s, CPU oracle (C restatement of LbmTools.h, {how})"}


ERROR_HORIZON = 600   # steps from rest for the velocity error (100 warm-up + 500 timed)


def velocity_error(nx, N, storage, compute, ulb, omega, steps=ERROR_HORIZON):
    """Velocity of the bench's storage/compute against a binary64 run of the
    same cavity, both `steps` steps from rest (independent of --steps)."""
    lat = nx.DenseLattice((N, N, N), storage, compute)
    lat.setup_cavity(ulb)
    lat.run(steps, omega)
    _, u = lat.moments()
    del lat
    ref = nx.DenseLattice((N, N, N), "f64", "f64")
    ref.setup_cavity(ulb)
    ref.run(steps, omega)
    _, u64 = ref.moments()
    del ref
    d = u.astype(np.float64) - u64
    return {"steps": steps, "max_rel": float(np.abs(d).max() / np.abs(u64).max()),
            "l2_rel": float(np.sqrt((d * d).sum() / (u64 * u64).sum())),
            "norms": "max |u - u64| / max |u64| and ||u - u64||_2 / ||u64||_2 over all cells",
            "reference": "binary64 GPU run of the same case (bit-exact with the CPU oracle)"}


def batch_velocity_error(nx, members, storage, steps=ERROR_HORIZON, every=8):
    """Velocity of the batch in `storage` against a binary64 batch of the same
    members, both `steps` steps from rest, on every `every`-th member: the
    largest max |u - u64| / max |u64| of a member and the relative L2 norm
    over the sampled members' cells."""
    out = {}
    for st in (storage, "f64"):
        b = nx.CavityBatch(members, st)
        b.init()
        b.step(steps)
        out[st] = [b.moments(k)[1].astype(np.float64) for k in range(0, len(members), every)]
        del b
    per, d2, n2, skipped = [], 0.0, 0.0, 0
    for u, u64 in zip(out[storage], out["f64"]):
        if not (np.isfinite(u).all() and np.isfinite(u64).all()) or max(np.abs(u).max(), np.abs(u64).max()) > 1.0:
            skipped += 1           # diverging: BGK unstable at this member's Re (tau near 1/2)
            continue
        d = u - u64
        per.append(float(np.abs(d).max() / max(np.abs(u64).max(), 1e-300)))
        d2 += float((d * d).sum())
        n2 += float((u64 * u64).sum())
    return {"steps": steps, "members_sampled": len(out["f64"]), "members_diverging": skipped,
            "max_rel": max(per), "median_member_max_rel": float(np.median(per)),
            "l2_rel": float(np.sqrt(d2 / n2)),
            "norms": "per-member max |u - u64| / max |u64| (largest, median); ||u - u64||_2 / ||u64||_2 over "
                     "the sampled members' cells; members diverging in either run (not finite, or |u| > 1 "
                     "lattice units: 20 times the lid speed) left out",
            "reference": "binary64 batch of the same members (bit-exact with the CPU oracle)"}


def compute_ref_rate(nx, N, storage, ulb, omega, steps, stream):
    """MLUPS of the same container step with the reference's float-path
    collision (NX_COMPUTE_REF: LbmTools.h:209-281 with its double literals)."""
    lat = nx.DenseLattice((N, N, N), storage, "ref")
    lat.setup_cavity(ulb, stream=stream.handle)
    lat.run(20, omega, stream=stream.handle)
    stream.sync()
    e0, e1 = nx.Event(), nx.Event()
    e0.record(stream)
    lat.run(steps, omega, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1)
    return {"value": round((N - 2) ** 3 * steps / (ms * 1e-3) / 1e6, 1), "unit": "MLUPS", "steps": steps,
            "ms_per_step": round(ms / steps, 5), "compute": "ref (binary32 with the reference's binary64 literals)"}


def copy_calibration(nx):
    """Plain device copy of 1 GiB, 10 launches (nx_copy_kernel): what
    FETCH_SIZE / WRITE_SIZE read for a kernel that moves a known number of
    bytes (1 GiB read + 1 GiB written per launch)."""
    from neon_amd import abi
    n = 1 << 30
    a, b = nx.DeviceBuffer(n), nx.DeviceBuffer(n)
    a.zero()
    s = nx.Stream()
    e0, e1 = nx.Event(), nx.Event()
    e0.record(s)
    for _ in range(10):
        abi.check(abi.lib.nx_copy_kernel(b.ptr, a.ptr, n, s.handle))
    e1.record(s)
    e1.sync()
    ms = e0.elapsed_ms(e1) / 10
    print(json.dumps({"copy_bytes_per_launch": 2 * n, "ms": ms, "GBs": 2 * n / ms / 1e6}))


def shifted_copy_calibration(nx, N=256, reps=20):
    """The dense step's access pattern without arithmetic (nx_shifted_copy:
    19 planes read at the upstream cells, 19 written) on an N^3 grid for 2-,
    4- and 8-byte elements: the ceiling of the pull-streaming step on SoA
    planes, beside the plain 1 GiB copy."""
    import ctypes as C
    from neon_amd import abi
    s = nx.Stream()
    out = {"grid": [N, N, N]}
    for eb in (2, 4, 8):
        nbytes = 19 * N ** 3 * eb
        a, b = nx.DeviceBuffer(nbytes), nx.DeviceBuffer(nbytes)
        a.zero()
        dim = (C.c_int32 * 3)(N, N, N)
        for _ in range(3):
            abi.check(abi.lib.nx_shifted_copy(b.ptr, a.ptr, dim, eb, s.handle))
        e0, e1 = nx.Event(), nx.Event()
        e0.record(s)
        for _ in range(reps):
            abi.check(abi.lib.nx_shifted_copy(b.ptr, a.ptr, dim, eb, s.handle))
        e1.record(s)
        e1.sync()
        ms = e0.elapsed_ms(e1) / reps
        out[f"{eb}B"] = {"us": round(ms * 1e3, 1), "GBs": round(2 * nbytes / ms / 1e6, 1),
                         "frac_of_8TBs": round(2 * nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}
        del a, b
    n = 1 << 30
    a, b = nx.DeviceBuffer(n), nx.DeviceBuffer(n)
    a.zero()
    for _ in range(3):
        abi.check(abi.lib.nx_copy_kernel(b.ptr, a.ptr, n, s.handle))
    e0, e1 = nx.Event(), nx.Event()
    e0.record(s)
    for _ in range(reps):
        abi.check(abi.lib.nx_copy_kernel(b.ptr, a.ptr, n, s.handle))
    e1.record(s)
    e1.sync()
    ms = e0.elapsed_ms(e1) / reps
    out["plain_copy_1GiB"] = {"us": round(ms * 1e3, 1), "GBs": round(2 * n / ms / 1e6, 1)}
    print(json.dumps({"shifted_copy_calibration": out}), flush=True)


def porous(nx, a):
    N, storage, compute, dsteps, dwarm = POROUS[a.workload][:5]
    variant = POROUS[a.workload][5] if len(POROUS[a.workload]) > 5 else None
    tiled = variant == "tiled"
    from neon_amd import abi as _abi
    _abi.check(_abi.lib.nx_d3q19_set_block_pairs(1 if variant == "pairs" else 0))
    steps = dsteps if a.steps is None else a.steps
    warmup = dwarm if a.warmup is None else a.warmup
    nu, g = 0.02, (1e-6, 0.0, 0.0)
    omega = 1.0 / (3.0 * nu + 0.5)
    nx.init(0)
    dim = (N, N, N)
    t0 = time.perf_counter()
    solid = nx.spheres(dim, 2000, 6.0, 20.0, seed=2024, periodic=1)
    grid = nx.BlockGrid(dim, 1 - solid, periodic=1)
    t_setup = time.perf_counter() - t0
    nb3 = (N // 8) ** 3
    stream = nx.Stream()

    def make(st, cm):
        lat = nx.BlockLattice(grid, st, cm)
        from neon_amd import abi
        act = np.unpackbits(grid.masks.view(np.uint8), bitorder="little").astype(bool)
        lat.set_classes(np.where(act, abi.NX_CELL_BULK, abi.NX_CELL_UNDEFINED).astype(np.uint32),
                        stream=stream.handle)
        return lat

    lat = make(storage, compute)
    if tiled:
        lat.set_tiled(True)
    lat.run(warmup, omega, force=g, stream=stream.handle)
    stream.sync()
    e0, e1 = nx.Event(), nx.Event()
    nx.sync()
    tw = time.perf_counter()
    e0.record(stream)
    lat.run(steps, omega, force=g, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    wall = time.perf_counter() - tw
    ms = e0.elapsed_ms(e1)
    ms_step = ms / steps
    fluid = grid.n_active
    mlups = fluid * steps / (ms * 1e-3) / 1e6
    b_alg = 2 * 19 * STORAGE_BYTES[storage]
    achieved = fluid * b_alg / (ms_step * 1e-3) / 1e9
    err = None
    if not a.no_error:
        # a fixed horizon from rest, whatever --steps and --warmup are
        del lat
        lat = make(storage, compute)
        if tiled:
            lat.set_tiled(True)
        lat.run(POROUS_ERROR_STEPS, omega, force=g)
        _, u = lat.moments()
        del lat
        ref = make("f64", "f64")
        ref.run(POROUS_ERROR_STEPS, omega, force=g)
        _, u64 = ref.moments()
        act = np.repeat(np.unpackbits(grid.masks.view(np.uint8), bitorder="little")
                        .astype(bool).reshape(-1, 512), 3, axis=0).reshape(-1)
        d = (u.astype(np.float64) - u64)[act]
        err = {"steps": POROUS_ERROR_STEPS, "max_rel": float(np.abs(d).max() / np.abs(u64[act]).max()),
               "l2_rel": float(np.sqrt((d * d).sum() / (u64[act] ** 2).sum())),
               "reference": "binary64 GPU run of the same case"}
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": storage,
        "data": "synthetic (seeded sphere pack, fluid at rest, f = w_q)",
        "config": {"workload": f"porous medium D3Q19 BGK on bGrid {N}^3 (8^3 blocks), 2000 "
                               f"spheres r~U[6,20) seed 2024, periodic x, g=1e-6 x, nu=0.02, "
                               f"{storage} storage + {compute} collision, {steps} timed steps "
                               f"after {warmup} warm-up",
                   "grid": "bGrid", "N": N, "blocks": grid.n_blocks, "candidate_blocks": nb3,
                   "step": ("LDS-staged, one workgroup per 2x2x1 tile of blocks" if tiled
                            else "one workgroup per block, two voxels per lane" if variant == "pairs"
                            else "one workgroup per block"),
                   "active_block_fraction": round(grid.n_blocks / nb3, 4),
                   "fluid_voxels": fluid, "fluid_fraction": round(fluid / N ** 3, 4),
                   "setup_s": round(t_setup, 3), "device": nx.device_name()},
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(f"d3q19_block{'_' + variant if variant else ''}_{storage}_{N}"),
                     "kernel": f"{'k_step_tiled' if tiled else 'k_step_pair' if variant == 'pairs' else 'k_step'}"
                               f"<block,{storage},{compute}> "
                               f"({b_alg} B/fluid voxel)"},
        "velocity_error_vs_f64": err,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def zouhe_duct(nx, a):
    X, Y, storage, dsteps, dwarm = ZOUHE[a.workload]
    steps = dsteps if a.steps is None else a.steps
    warmup = dwarm if a.warmup is None else a.warmup
    from neon_amd import abi
    nu = 0.02
    omega = 1.0 / (3.0 * nu + 0.5)
    nx.init(0)
    dim = (X, Y, Y)
    cls = np.full((Y, Y, X), abi.NX_CELL_BULK, np.uint32)
    cls[:, 0, :] = cls[:, -1, :] = cls[0, :, :] = cls[-1, :, :] = abi.NX_CELL_BOUNCEBACK
    zh = np.zeros((Y, Y, X), np.uint32)
    zh[:, :, 0] = abi.NX_ZH_XMIN
    zh[:, :, -1] = abi.NX_ZH_XMAX
    zh[cls != abi.NX_CELL_BULK] = 0
    values = [1.0005, 0.9995, 0.0, 0.0, 0.0, 0.0]
    lat = nx.DenseLattice(dim, storage, storage)
    lat.set_classes(cls.reshape(-1), zouhe=zh.reshape(-1))
    stream = nx.Stream()
    lat.run_zouhe(warmup, omega, values, stream=stream.handle)
    stream.sync()
    e0, e1 = nx.Event(), nx.Event()
    nx.sync()
    tw = time.perf_counter()
    e0.record(stream)
    lat.run_zouhe(steps, omega, values, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    wall = time.perf_counter() - tw
    ms = e0.elapsed_ms(e1)
    ms_step = ms / steps
    fluid = int((cls == abi.NX_CELL_BULK).sum())
    mlups = fluid * steps / (ms * 1e-3) / 1e6
    b_alg = 2 * 19 * STORAGE_BYTES[storage]
    achieved = fluid * b_alg / (ms_step * 1e-3) / 1e9
    _, u = lat.moments(stream=stream.handle)
    ux = u[0].reshape(Y, Y, X).astype(np.float64)
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": storage,
        "data": "synthetic (duct at rest, f = w_q)",
        "config": {"workload": f"pressure-driven duct {X}x{Y}x{Y}, Zou-He density faces 1.0005 / 0.9995 "
                               f"on the x planes, bounce-back side walls, nu = 0.02, {storage}, "
                               f"{steps} timed steps after {warmup}",
                   "fluid_cells": fluid, "zouhe_cells": int((zh > 0).sum()), "device": nx.device_name(),
                   "ux_axis_x32": float(ux[Y // 2, Y // 2, 32])},
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(f"d3q19_zouhe_{storage}_{X}"),
                     "kernel": f"k_step_zh<{storage}> ({b_alg} B/fluid cell)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def batched(nx, a):
    n, storage, dsteps, dwarm = BATCH[a.workload]
    steps = dsteps if a.steps is None else a.steps
    warmup = dwarm if a.warmup is None else a.warmup
    steps += steps % 2                    # the batch steps in pairs (arena ping-pong)
    warmup += warmup % 2
    ulb = 0.05
    nx.init(0)
    t0 = time.perf_counter()
    members = nx.batch_cavities(n, seed=2025, ulb=ulb)
    fluid = sum(int((m["classes"] == 2).sum()) for m in members)
    b = nx.CavityBatch(members, storage)
    t_setup = time.perf_counter() - t0
    if a.edge is not None:
        b.set_variant(xcd_chunks=b.launch()["xcd_chunks"], edge=a.edge)
    if a.autotune:
        b.autotune()
    launch = b.launch()
    b.init()
    stream = nx.Stream()
    b.step(warmup, stream=stream.handle)
    stream.sync()
    e0, e1 = nx.Event(), nx.Event()
    nx.sync()
    tw = time.perf_counter()
    e0.record(stream)
    b.step(steps, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    wall = time.perf_counter() - tw
    ms = e0.elapsed_ms(e1)
    ms_step = ms / steps
    cells = b.n_cells
    mlups = fluid * steps / (ms * 1e-3) / 1e6
    b_alg = 2 * 19 * STORAGE_BYTES[storage]
    achieved = fluid * b_alg / (ms_step * 1e-3) / 1e9
    err = None
    if not a.no_error:
        err = batch_velocity_error(nx, members, storage)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(64, storage, "f32", 1.0 / (3.0 * ulb * 64 / 500.0 + 0.5), max_steps=4000)
        cpu["sample"] = "oracle step of one 64^3 member-sized cavity, " + cpu["sample"]
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": storage,
        "data": "synthetic (seeded cavities with box obstacles, initialised at rest)",
        "config": {"workload": f"{n} batched lid-driven cavities, edge U{{32..96}} step 8, "
                               f"Re U[50,1000], 0-4 boxes U{{4..16}}, {storage} storage, one "
                               f"persistent launch per step, {steps} timed steps after {warmup}",
                   "members": n, "cells": cells, "fluid_cells": fluid,
                   "arena_GB": round(2 * 19 * STORAGE_BYTES[storage] * cells / 1e9, 2),
                   "setup_s": round(t_setup, 2), "launch": launch, "device": nx.device_name()},
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(f"d3q19_batch_{storage}_{n}"),
                     "kernel": f"k_batch_step<{storage}> ({b_alg} B/fluid cell)"},
        "velocity_error_vs_f64": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def batch_converge(nx, a):
    n, storage, tol, every, cap = CONVERGE[a.workload]
    max_steps = cap if a.steps is None else a.steps
    max_steps -= max_steps % every
    chunk = 5000
    ulb = 0.05
    nx.init(0)
    t0 = time.perf_counter()
    members = nx.batch_cavities(n, seed=2025, ulb=ulb)
    fluid = np.array([int((m["classes"] == 2).sum()) for m in members])
    edge = np.array([m["dim"][0] for m in members])
    re = np.array([m["re"] for m in members])
    b = nx.CavityBatch(members, storage)
    b.init()
    t_setup = time.perf_counter() - t0
    m0 = np.array([b.mass(k) for k in range(n)])
    nx.sync()
    tw = time.perf_counter()
    done, left = 0, n
    while done < max_steps and left > 0:
        k = min(chunk, max_steps - done)
        steps, res, left = b.run(max_steps=k, check_every=every, tol=tol)
        done += k
        print(f"batch512_converge: {done} steps, {n - left} of {n} finished (converged or diverged), "
              f"{time.perf_counter() - tw:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - tw
    m1 = np.array([b.mass(k) for k in range(n)])
    conv = np.isfinite(res) & (res < tol)
    div = ~np.isfinite(res)              # the flow blew up: the member stopped stepping
    run_on = ~conv & ~div
    updates = float((fluid * steps).sum())
    bins = list(range(32, 97, 8))
    by_edge = {str(e): {"members": int((edge == e).sum()), "converged": int(conv[edge == e].sum()),
                        "diverged": int(div[edge == e].sum()),
                        "median_steps": int(np.median(steps[(edge == e) & conv])) if conv[edge == e].any() else None}
               for e in bins}
    re_bins = [(50, 200), (200, 400), (400, 600), (600, 800), (800, 1000)]
    by_re = {f"{lo}-{hi}": {"members": int(((re >= lo) & (re < hi)).sum()),
                            "converged": int(conv[(re >= lo) & (re < hi)].sum()),
                            "diverged": int(div[(re >= lo) & (re < hi)].sum())} for lo, hi in re_bins}
    # cell Reynolds number U dx / nu = Re / edge: BGK loses stability as tau -> 1/2
    cell_re = re / edge
    hist_edges = [0, 2000, 5000, 10000, 20000, 50000, 100000, 200000, 500000, 1000000]
    h, _ = np.histogram(steps[conv], bins=hist_edges)
    ok = np.isfinite(m1)
    drift = np.abs(m1[ok] - m0[ok]) / m0[ok]
    out = {
        "metric": METRIC, "value": round(updates / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": int(steps.max()), "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": storage,
        "data": "synthetic (seeded cavities with box obstacles, started at rest)",
        "config": {"workload": f"{n} batched lid-driven cavities, edge U{{32..96}} step 8, Re U[50,1000], "
                               f"0-4 boxes U{{4..16}}, {storage} storage, run to residual {tol:g} "
                               f"(K={every}) or {max_steps} steps, converged members retire",
                   "members": n, "fluid_cells": int(fluid.sum()), "setup_s": round(t_setup, 2),
                   "launch": b.launch(), "device": nx.device_name()},
        "wall_s": round(wall, 2),
        "converged": int(conv.sum()), "unconverged": int(n - conv.sum()),
        "diverged": int(div.sum()),
        "cell_reynolds": {"diverged_min": float(cell_re[div].min()) if div.any() else None,
                          "stable_max": float(cell_re[~div].max())},
        "step_cap": max_steps,
        "steps_histogram": {"bin_edges": hist_edges, "converged_members": h.tolist()},
        "by_edge": by_edge, "by_re": by_re,
        "residual_of_unconverged": {"min": float(res[run_on].min()) if run_on.any() else None,
                                    "median": float(np.median(res[run_on])) if run_on.any() else None,
                                    "max": float(res[run_on].max()) if run_on.any() else None},
        "mass_drift_rel": {"members": int(ok.sum()), "max": float(drift.max()),
                           "median": float(np.median(drift))},
        "cell_updates": updates,
    }
    print(json.dumps(out), flush=True)


def mres_cavity(nx, a):
    N, sdf, storage, dsteps, dwarm, re, *rest = MRES[a.workload]
    lattice = rest[0] if rest else "bgk19"
    nq = 27 if lattice == "kbc27" else 19
    steps = dsteps if a.steps is None else a.steps
    warmup = dwarm if a.warmup is None else a.warmup
    from neon_amd import mgrid as mg
    nx.init(0)
    sphere = N == "sphere"
    t0 = time.perf_counter()
    if sphere:
        scale = sdf
        dim, lams, cls, _ = mg.sphere_problem(scale)
        depth = 3
        ulb, om0 = mg.sphere_parameters(scale, re, storage)
        g = mg.MGrid(dim, list(lams), storage, lattice=lattice)
        g.setup_cells(cls, (ulb, 0.0, 0.0))
        problem = f"flow over a sphere (--scale {scale}) {dim[0]}x{dim[1]}x{dim[2]}, 3 levels, inlet u {ulb:.4g}"
    else:
        dim = (N, N, N)
        depth = len(sdf) - 1
        lams = mg.cavity_lambdas(dim, sdf)
        ulb, om0 = mg.cavity_parameters(dim, re, storage, depth=depth)
        g = mg.MGrid(dim, list(lams), storage, lattice=lattice)
        g.setup_cavity(ulb)
        problem = f"{depth}-level refined lid-driven cavity {N}^3 (levelSDF {list(sdf)}), ulb {ulb:.4g}"
    t_setup = time.perf_counter() - t0
    info = g.info()
    stream = nx.Stream()
    g.run(warmup, om0, stream=stream.handle)
    stream.sync()
    e0, e1 = nx.Event(), nx.Event()
    nx.sync()
    tw = time.perf_counter()
    e0.record(stream)
    g.run(steps, om0, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    wall = time.perf_counter() - tw
    ms = e0.elapsed_ms(e1)
    ms_step = ms / steps
    leaf = info["leaf"]
    # the reference's count (lbmMultiRes.h:385-391): level l's leaves updated
    # 2^(depth-1-l) times per coarse step
    updates = sum(n << (depth - 1 - l) for l, n in enumerate(leaf))
    mlups = updates * steps / (ms * 1e-3) / 1e6
    emlups = (1 << (depth - 1)) * dim[0] * dim[1] * dim[2] * steps / (ms * 1e-3) / 1e6
    s_b = STORAGE_BYTES[storage]
    # bytes the kernels must move per coarse step: every pass reads and
    # writes the 19 populations of a leaf; a sub-step of level l is one pass
    # when the level's stream runs fused with its next collide (steady
    # state inside a call), else two (collide, stream)
    passes = [(1 if fused else 2) << (depth - 1 - l) for l, fused in enumerate(info["fused_levels"])]
    b_step = sum(n * p for n, p in zip(leaf, passes)) * 2 * nq * s_b
    achieved = b_step / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": storage,
        "data": f"synthetic ({'flow over a sphere' if sphere else f'{depth}-level refined lid-driven cavity'}"
                f" started at rest)",
        "config": {"workload": f"{problem}, {'D3Q27 KBC' if nq == 27 else 'D3Q19 BGK'}, "
                               f"{storage}, Re {re:g}, {steps} timed coarse steps after {warmup}",
                   "leaves": leaf, "active": info["active"], "blocks": info["blocks"],
                   "store_links": info["store_links"], "setup_s": round(t_setup, 2), "omega0": om0,
                   "device": nx.device_name()},
        "mlups_effective": round(emlups, 1),
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"coarse step: k_mr_collide + k_mr_stream<{storage}, {lattice}>, passes per level "
                               f"{passes}, 2*{nq}*{s_b} B per leaf per pass, {b_step / 1e9:.3f} GB"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="cavity256_f32",
                    choices=sorted(WORKLOADS) + sorted(POROUS) + sorted(BATCH) + sorted(CONVERGE) + sorted(ZOUHE)
                    + sorted(MRES))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-error", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--autotune", action="store_true",
                    help="time every launch configuration first and keep the fastest")
    ap.add_argument("--vx", type=int, default=None,
                    help="cells per lane of the dense step (1, or 16 bytes of cells: 4 f32, 2 f64)")
    ap.add_argument("--boundary", type=int, default=None,
                    help="dense step boundary handling: 0 interior bitmap + full pull, 1 one collision with fix-ups, "
                         "2 as 1 with the interior bitmap, 3 as 2 with edge speculation")
    ap.add_argument("--edge", type=int, default=None, choices=(0, 1),
                    help="batched cavities: 1 edge-speculating step, 0 wall links fetched after the flags")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel (1 GiB, 10 launches): the "
                         "reference for scaling FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--shifted", action="store_true",
                    help="with --copy-calibration: time the step's 19-plane shifted copy (no arithmetic) "
                         "at 256^3 for 2/4/8-byte elements, the ceiling of the pull-streaming step")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU z-slab decomposition is a design, not implemented)",
              file=sys.stderr)
        sys.exit(2)

    import neon_amd as nx

    if a.copy_calibration:
        if a.shifted:
            nx.init(0)
            shifted_copy_calibration(nx)
        else:
            copy_calibration(nx)
        return

    if a.workload in POROUS:
        return porous(nx, a)
    if a.workload in BATCH:
        return batched(nx, a)
    if a.workload in CONVERGE:
        return batch_converge(nx, a)
    if a.workload in ZOUHE:
        return zouhe_duct(nx, a)
    if a.workload in MRES:
        return mres_cavity(nx, a)

    N, storage, compute, re, dsteps, dwarm = WORKLOADS[a.workload]
    steps = dsteps if a.steps is None else a.steps
    warmup = dwarm if a.warmup is None else a.warmup
    ulb = 0.05
    nu = ulb * N / re
    omega = 1.0 / (3.0 * nu + 0.5)

    nx.init(0)
    stream = nx.Stream()
    lat = nx.DenseLattice((N, N, N), storage, compute)
    lat.setup_cavity(ulb, stream=stream.handle)
    if a.vx is not None or a.boundary is not None:
        t = lat.tuning()
        if a.vx is not None:
            t["vx"] = a.vx
        if a.boundary is not None:
            t["boundary"] = a.boundary
        lat.set_tuning(**t)
    if a.autotune:
        lat.autotune(omega, stream=stream.handle)
        lat.setup_cavity(ulb, stream=stream.handle)
    lat.run(warmup, omega, stream=stream.handle)
    stream.sync()

    e0, e1 = nx.Event(), nx.Event()
    nx.sync()
    t0 = time.perf_counter()
    e0.record(stream)
    lat.run(steps, omega, stream=stream.handle)
    e1.record(stream)
    e1.sync()
    nx.sync()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    ms_step = ms / steps

    fluid = (N - 2) ** 3
    mlups = fluid * steps / (ms * 1e-3) / 1e6
    mlups_ref = N ** 3 * steps / (ms * 1e-3) / 1e6
    b_alg = 2 * 19 * STORAGE_BYTES[storage]
    achieved = fluid * b_alg / (ms_step * 1e-3) / 1e9
    tag = f"d3q19_dense_{storage}_{N}"
    traffic = traffic_from_profiles(tag)

    launch = lat.tuning()
    del lat
    ref_rate = None
    if storage == "f32" and compute == "f32":
        ref_rate = compute_ref_rate(nx, N, storage, ulb, omega, min(steps, 200), stream)

    err = None
    if not a.no_error and storage != "f64":
        err = velocity_error(nx, N, storage, compute, ulb, omega)

    cpu = cpu_all = None
    if not a.no_cpu:
        cpu = cpu_baseline(N, storage, compute, omega, threads=1, max_steps=40)
        cpu_all = cpu_baseline(N, storage, compute, omega)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": storage,
        "data": "synthetic (lid-driven cavity initialised at rest, f = w_q)",
        "config": {"workload": f"lid-driven cavity D3Q19 BGK, dense {N}^3 ({N-2}^3 fluid), "
                               f"{storage} storage + {compute} collision, Re={re:g}, "
                               f"lid u=({ulb},0,0), {steps} timed steps after {warmup} warm-up",
                   "grid": "dGrid", "N": N, "storage": storage, "compute": compute,
                   "omega": omega, "device": nx.device_name(), "launch": launch},
        "mlups_ref_convention": round(mlups_ref, 1),
        "mlups_compute_ref": ref_rate,
        "wall_s": round(wall, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": f"{dense_kernel_name(launch)}<{storage},{compute}> ({b_alg} B/fluid cell)"},
        "velocity_error_vs_f64": err,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
