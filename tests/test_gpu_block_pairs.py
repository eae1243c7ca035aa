"""The block-sparse step at 2-byte storage with two x-adjacent voxels per
lane (nx_d3q19_set_block_pairs(1): k_step_pair) against the one-voxel-per-lane
step and the CPU oracle, bit for bit, whole population arrays: the cavity on
grids with partial blocks, the porous periodic forced case with NX_FLAG_MIXED
voxels, and the 384^3 porous benchmark grid."""
import numpy as np
import pytest

from test_gpu_block import _raw, spheres

pytestmark = pytest.mark.gpu


@pytest.fixture
def pairs(nx):
    from neon_amd import abi

    def set_pairs(on):
        abi.check(abi.lib.nx_d3q19_set_block_pairs(1 if on else 0))
    yield set_pairs
    set_pairs(False)


def _run_both(nx, grid, setup, steps, omega, force, pairs):
    out = []
    for on in (False, True):
        pairs(on)
        lat = nx.BlockLattice(grid, "f16dev", "f32")
        setup(lat)
        lat.run(steps, omega, force=force)
        out.append(_raw(lat.populations(), lat.storage))
        del lat
    return out


def test_pairs_cavity_bit_exact(nx, orc, pairs):
    dim = (40, 36, 17)
    print(f"pairs cavity {dim}", flush=True)
    act = np.ones(dim[0] * dim[1] * dim[2], np.uint8)
    grid = nx.BlockGrid(dim, act)
    one, two = _run_both(nx, grid, lambda l: l.setup_cavity(0.05), 13, 1.7, (0.0, 0.0, 0.0), pairs)
    assert np.array_equal(one, two)
    o = orc.BlockOracle(dim, act, 3, 1)
    o.setup_cavity(0.05)
    o.run(13, 1.7)
    keep = o.active_slots().reshape(-1, 512)
    p3, po = two.reshape(-1, 19, 512), _raw(o.fin, 3).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, keep], po.transpose(1, 0, 2)[:, keep])


def test_pairs_porous_periodic_forced_bit_exact(nx, orc, pairs):
    dim = (40, 24, 32)
    print(f"pairs porous {dim}", flush=True)
    solid = spheres(dim, 12, 3, 7, seed=5)
    solid[:8, :8, :] = True
    act = (~solid).astype(np.uint8).reshape(-1)
    grid = nx.BlockGrid(dim, act, periodic=1)
    keep = np.unpackbits(grid.masks.view(np.uint8), bitorder="little").astype(bool)
    cls = np.where(keep, 2, 3).astype(np.uint32)
    rng = np.random.default_rng(9)
    cls[keep & (rng.random(keep.size) < 0.03)] = 0
    one, two = _run_both(nx, grid, lambda l: l.set_classes(cls), 11, 1.6, (3e-5, 0.0, 0.0), pairs)
    assert np.array_equal(one, two)
    o = orc.BlockOracle(dim, act, 3, 1, periodic=1)
    o.set_classes(cls)
    o.run(11, 1.6, force=(3e-5, 0.0, 0.0))
    k3 = keep.reshape(-1, 512)
    p3, po = two.reshape(-1, 19, 512), _raw(o.fin, 3).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])


@pytest.mark.parametrize("dim", [(24, 40, 24), (8, 8, 16), (8, 24, 8)])
def test_pairs_periodic_every_axis_bit_exact(nx, orc, pairs, dim):
    """Periodic in x, y and z, odd block counts and one-block-wide domains
    (the pair at a block's x face reading its own block across the wrap)."""
    print(f"pairs periodic xyz {dim}", flush=True)
    n = dim[0] * dim[1] * dim[2]
    solid = spheres(dim, 3, 2, 3, seed=11)
    act = (~solid).astype(np.uint8).reshape(-1) if n > 4096 else np.ones(n, np.uint8)
    grid = nx.BlockGrid(dim, act, periodic=7)
    keep = np.unpackbits(grid.masks.view(np.uint8), bitorder="little").astype(bool)
    cls = np.where(keep, 2, 3).astype(np.uint32)
    one, two = _run_both(nx, grid, lambda l: l.set_classes(cls), 9, 1.6, (2e-5, 1e-5, 0.0), pairs)
    assert np.array_equal(one, two)
    o = orc.BlockOracle(dim, act, 3, 1, periodic=7)
    o.set_classes(cls)
    o.run(9, 1.6, force=(2e-5, 1e-5, 0.0))
    k3 = keep.reshape(-1, 512)
    p3, po = two.reshape(-1, 19, 512), _raw(o.fin, 3).reshape(-1, 19, 512)
    assert np.array_equal(p3.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])


def test_pairs_equal_single_on_the_benchmark_grid(nx, pairs):
    from neon_amd import abi
    dim = (384, 384, 384)
    print("pairs porous 384^3", flush=True)
    solid = nx.spheres(dim, 2000, 6.0, 20.0, seed=2024, periodic=1)
    grid = nx.BlockGrid(dim, 1 - solid, periodic=1)
    act = np.unpackbits(grid.masks.view(np.uint8), bitorder="little").astype(bool)
    cls = np.where(act, abi.NX_CELL_BULK, abi.NX_CELL_UNDEFINED).astype(np.uint32)
    one, two = _run_both(nx, grid, lambda l: l.set_classes(cls), 6, 1.0 / (3 * 0.02 + 0.5),
                         (1e-6, 0.0, 0.0), pairs)
    assert np.array_equal(one, two)
