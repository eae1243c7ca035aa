"""The block-sparse step with the block neighbourhood staged in LDS
(nx_d3q19_step_block_tiled: one workgroup per 2 x 2 x 1 tile of 8^3 blocks)
against the CPU oracle and against the per-block step, bit for bit: the
cavity on grids whose block counts are odd (tiles with absent members), the
porous periodic forced case with NX_FLAG_MIXED voxels, and the 384^3 porous
benchmark grid (staged and per-block step, whole population arrays)."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_block import _raw, spheres

pytestmark = pytest.mark.gpu


def _tile_count(origins):
    o = origins.reshape(-1, 3)
    return len({(x >> 4, y >> 4, z >> 3) for x, y, z in o.tolist()})


@pytest.mark.parametrize("st,cm,sto,cmo", [("f32", "f32", 1, 1), ("f32", "ref", 1, 2),
                                           ("f32dev", "f32", 2, 1), ("f16dev", "f32", 3, 1)])
def test_tiled_cavity_bit_exact(nx, orc, st, cm, sto, cmo):
    dim = (40, 36, 17)      # 5 x 5 x 3 blocks: tiles with one, two or four members
    print(f"tiled cavity {dim} {st}/{cm}", flush=True)
    act = np.ones(dim[0] * dim[1] * dim[2], np.uint8)
    grid = nx.BlockGrid(dim, act)
    g = nx.BlockLattice(grid, st, cm)
    ref = nx.BlockLattice(grid, st, cm)
    o = orc.BlockOracle(dim, act, sto, cmo)
    g.set_tiled(True)
    assert g.n_tiles == _tile_count(grid.origins) == 3 * 3 * 3
    for lat in (g, ref, o):
        lat.setup_cavity(0.05)
    for lat in (g, ref, o):
        lat.run(13, 1.7)
    pg = _raw(g.populations(), g.storage)
    assert np.array_equal(pg, _raw(ref.populations(), ref.storage))   # every slot
    keep = o.active_slots().reshape(-1, 512)
    p3, po = pg.reshape(-1, 19, 512), _raw(o.fin, sto).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, keep], po.transpose(1, 0, 2)[:, keep])


@pytest.mark.parametrize("st,sto", [("f32", 1), ("f16dev", 3), ("f32dev", 2)])
def test_tiled_porous_periodic_forced_bit_exact(nx, orc, st, sto):
    dim = (40, 24, 32)
    print(f"tiled porous {dim} {st}", flush=True)
    solid = spheres(dim, 12, 3, 7, seed=5)
    solid[:8, :8, :] = True           # absent blocks
    act = (~solid).astype(np.uint8).reshape(-1)
    grid = nx.BlockGrid(dim, act, periodic=1)
    g = nx.BlockLattice(grid, st, "f32")
    ref = nx.BlockLattice(grid, st, "f32")
    o = orc.BlockOracle(dim, act, sto, 1, periodic=1)
    g.set_tiled(True)
    keep = o.active_slots()
    cls = np.where(keep, 2, 3).astype(np.uint32)
    rng = np.random.default_rng(9)
    cls[keep & (rng.random(keep.size) < 0.03)] = 0
    for lat in (g, ref, o):
        lat.set_classes(cls)
    fl = g.get_flags()
    assert ((fl[keep] & (1 << 29)) != 0).sum() > 10      # NX_FLAG_MIXED voxels
    for lat in (g, ref, o):
        lat.run(11, 1.6, force=(3e-5, 0.0, 0.0))
    pg = _raw(g.populations(), g.storage)
    assert np.array_equal(pg, _raw(ref.populations(), ref.storage))
    k3 = keep.reshape(-1, 512)
    p3, po = pg.reshape(-1, 19, 512), _raw(o.fin, sto).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])


@pytest.mark.parametrize("dim", [(24, 40, 24), (8, 8, 16), (8, 24, 8)])
@pytest.mark.parametrize("st,sto", [("f32", 1), ("f16dev", 3)])
def test_tiled_periodic_every_axis_bit_exact(nx, orc, dim, st, sto):
    """Periodic in x, y and z on odd block counts and one-block-wide domains
    (a block its own neighbour): tiles whose absent member's extent holds
    wrapped upstream voxels, read through the connectivity rows."""
    print(f"tiled periodic xyz {dim} {st}", flush=True)
    n = dim[0] * dim[1] * dim[2]
    solid = spheres(dim, 3, 2, 3, seed=11)
    act = (~solid).astype(np.uint8).reshape(-1) if n > 4096 else np.ones(n, np.uint8)
    grid = nx.BlockGrid(dim, act, periodic=7)
    g = nx.BlockLattice(grid, st, "f32")
    ref = nx.BlockLattice(grid, st, "f32")
    o = orc.BlockOracle(dim, act, sto, 1, periodic=7)
    g.set_tiled(True)
    keep = o.active_slots()
    cls = np.where(keep, 2, 3).astype(np.uint32)
    for lat in (g, ref, o):
        lat.set_classes(cls)
    for lat in (g, ref, o):
        lat.run(9, 1.6, force=(2e-5, 1e-5, 0.0))
    pg = _raw(g.populations(), g.storage)
    assert np.array_equal(pg, _raw(ref.populations(), ref.storage))
    k3 = keep.reshape(-1, 512)
    p3, po = pg.reshape(-1, 19, 512), _raw(o.fin, sto).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])


@pytest.mark.parametrize("st", ["f16dev", "f32"])
def test_tiled_equals_per_block_step_on_the_benchmark_grid(nx, st):
    """384^3, 2000 spheres r~U[6,20) (seed 2024), periodic x, g = 1e-6: the
    porous benchmark's grid; 6 steps of each step from the same start."""
    from neon_amd import abi
    dim = (384, 384, 384)
    print(f"tiled porous 384^3 {st}", flush=True)
    solid = nx.spheres(dim, 2000, 6.0, 20.0, seed=2024, periodic=1)
    grid = nx.BlockGrid(dim, 1 - solid, periodic=1)
    act = np.unpackbits(grid.masks.view(np.uint8), bitorder="little").astype(bool)
    cls = np.where(act, abi.NX_CELL_BULK, abi.NX_CELL_UNDEFINED).astype(np.uint32)
    lats = []
    for tiled in (False, True):
        lat = nx.BlockLattice(grid, st, "f32")
        lat.set_classes(cls)
        if tiled:
            lat.set_tiled(True)
        lat.run(6, 1.0 / (3 * 0.02 + 0.5), force=(1e-6, 0.0, 0.0))
        lats.append(_raw(lat.populations(), lat.storage))
        del lat
    assert np.array_equal(lats[0], lats[1])


def test_tiled_rejects_what_it_cannot_run(nx):
    from neon_amd import abi
    dim = (16, 16, 16)
    act = np.ones(16 ** 3, np.uint8)
    grid = nx.BlockGrid(dim, act)
    g64 = nx.BlockLattice(grid, "f64", "f64")
    g64.setup_cavity(0.05)
    g64.set_tiled(True)
    with pytest.raises(abi.NeonError, match="2- and 4-byte"):
        g64.run(1, 1.5)
    other = nx.BlockGrid(dim, act)      # the same block count, other tables
    g32 = nx.BlockLattice(other, "f32", "f32")
    g32.setup_cavity(0.05)
    p = nx.params(1.5)
    rc = abi.lib.nx_d3q19_step_block_tiled(C.byref(g32.desc), g64._tiles, g32.fin.ptr, g32.fout.ptr,
                                           g32.flags.ptr, C.byref(p), None)
    assert rc == abi.NX_ERR_INVALID
