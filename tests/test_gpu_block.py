"""GPU parity of the block-sparse (bGrid, 8^3 blocks) D3Q19 path against the
CPU oracle, bit for bit, on the cavity and on a porous periodic forced case."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMBOS = [("f64", "f64", 0, 0), ("f32", "f32", 1, 1), ("f32", "ref", 1, 2),
          ("f32dev", "f32", 2, 1), ("f16dev", "f32", 3, 1)]


def _raw(a, storage):
    return a.view(np.uint16) if storage == 3 else a


def spheres(dim, n, rmin, rmax, seed):
    X, Y, Z = dim
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    solid = np.zeros((Z, Y, X), bool)
    for _ in range(n):
        c = rng.random(3) * (X, Y, Z)
        r = rng.uniform(rmin, rmax)
        solid |= (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r
    return solid


@pytest.mark.parametrize("st,cm,sto,cmo", COMBOS)
def test_block_cavity_bit_exact(nx, orc, st, cm, sto, cmo):
    dim = (24, 20, 17)          # not a multiple of 8 along y, z: partial blocks
    act = np.ones(dim[0] * dim[1] * dim[2], np.uint8)
    grid = nx.BlockGrid(dim, act)
    g = nx.BlockLattice(grid, st, cm)
    o = orc.BlockOracle(dim, act, sto, cmo)
    assert grid.n_blocks == o.n_blocks
    assert np.array_equal(grid.conn, o.conn) and np.array_equal(grid.masks, o.masks)
    g.setup_cavity(0.05)
    o.setup_cavity(0.05)
    keep = o.active_slots()
    assert np.array_equal(g.get_flags()[keep], o.flags[keep])
    g.run(13, 1.7)
    o.run(13, 1.7)
    pg = _raw(g.populations(), g.storage).reshape(-1, 19, 512)
    po = _raw(o.fin, sto).reshape(-1, 19, 512)
    k3 = keep.reshape(-1, 512)
    assert np.array_equal(pg.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])
    rg, ug = g.moments()
    ro, uo = o.moments()
    assert np.array_equal(rg[keep], ro[keep])
    u3g, u3o = ug.reshape(-1, 3, 512), uo.reshape(-1, 3, 512)
    assert np.array_equal(u3g.transpose(1, 0, 2)[:, k3], u3o.transpose(1, 0, 2)[:, k3])
    assert abs(g.mass() - o.mass()) <= 1e-12 * abs(o.mass())


@pytest.mark.parametrize("st,sto", [("f32", 1), ("f16dev", 3), ("f64", 0)])
def test_block_porous_periodic_forced_bit_exact(nx, orc, st, sto):
    dim = (40, 24, 32)
    solid = spheres(dim, 12, 3, 7, seed=5)
    solid[:8, :8, :] = True           # absent blocks
    act = (~solid).astype(np.uint8).reshape(-1)
    cm, cmo = ("f64", 0) if st == "f64" else ("f32", 1)
    grid = nx.BlockGrid(dim, act, periodic=1)
    g = nx.BlockLattice(grid, st, cm)
    o = orc.BlockOracle(dim, act, sto, cmo, periodic=1)
    assert grid.n_blocks == o.n_blocks and np.array_equal(grid.conn, o.conn)
    keep = o.active_slots()
    cls = np.where(keep, 2, 3).astype(np.uint32)
    # a few active bounce-back voxels next to inactive ones: voxels with both
    # kinds of wall link (NX_FLAG_MIXED)
    rng = np.random.default_rng(9)
    cls[keep & (rng.random(keep.size) < 0.03)] = 0
    g.set_classes(cls)
    o.set_classes(cls)
    fl = g.get_flags()
    assert np.array_equal(fl[keep], o.flags[keep])
    assert ((fl[keep] & (1 << 28)) != 0).sum() > 100
    assert ((fl[keep] & (1 << 29)) != 0).sum() > 10
    g.run(11, 1.6, force=(3e-5, 0.0, 0.0))
    o.run(11, 1.6, force=(3e-5, 0.0, 0.0))
    pg = _raw(g.populations(), g.storage).reshape(-1, 19, 512)
    po = _raw(o.fin, sto).reshape(-1, 19, 512)
    k3 = keep.reshape(-1, 512)
    assert np.array_equal(pg.transpose(1, 0, 2)[:, k3], po.transpose(1, 0, 2)[:, k3])


def test_block_residual(nx, orc):
    dim = (16, 16, 16)
    act = np.ones(16 ** 3, np.uint8)
    grid = nx.BlockGrid(dim, act)
    g = nx.BlockLattice(grid, "f32", "f32")
    g.setup_cavity(0.05)
    up = nx.DeviceBuffer(3 * grid.n_blocks * 512 * 4)
    up.zero()
    g.run(20, 1.5)
    a0, b0 = g.residual(up)
    assert a0 == pytest.approx(b0, rel=1e-12)        # from u_prev = 0
    a1, b1 = g.residual(up)
    assert a1 == 0.0 and b1 == pytest.approx(b0, rel=1e-12)   # unchanged field
    g.run(1, 1.5)
    a2, _ = g.residual(up)
    assert 0 < a2 < a0
