"""Block-sparse grid (bGrid, 8^3 blocks) structure and oracle checks on the CPU.

* the oracle builder against a brute-force numpy restatement, on the
  reference test's odd domain (17, 33, 71) (libNeonDomain/tests/
  domain-neighbour-globalIdx/src/runHelper.h:35),
* the neighbour-global-index property of the reference's domain tests
  (domain-neighbour-globalIdx/src/testsAndContainers.cu:60-105): following
  the connectivity from every active voxel in each of the 26 directions lands
  on global + direction, or on nothing when that voxel is inactive/outside,
* the library's host builder (nx_bgrid_create, nx_bgrid_create_cb) bit-exact
  against the oracle builder,
* the oracle's block step equal bit for bit to its dense step on the same
  geometry (a different addressing path, the same arithmetic).
"""
import ctypes as C

import numpy as np
import pytest


def brute_blocks(dim, active):
    X, Y, Z = dim
    act = active.reshape(Z, Y, X)
    nbx, nby, nbz = (X + 7) // 8, (Y + 7) // 8, (Z + 7) // 8
    order = []
    for bz in range(nbz):
        for by in range(nby):
            for bx in range(nbx):
                sub = act[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
                if sub.any():
                    order.append((bx, by, bz))
    rank = {b: i for i, b in enumerate(order)}
    origins = np.array([[8 * bx, 8 * by, 8 * bz] for bx, by, bz in order], np.int32).reshape(-1)
    masks = np.zeros((len(order), 16), np.uint32)
    conn = np.full((len(order), 27), 0xFFFFFFFF, np.uint32)
    for i, (bx, by, bz) in enumerate(order):
        for v in range(512):
            x, y, z = bx * 8 + (v & 7), by * 8 + ((v >> 3) & 7), bz * 8 + (v >> 6)
            if x < X and y < Y and z < Z and act[z, y, x]:
                masks[i, v >> 5] |= np.uint32(1 << (v & 31))
        for k in range(3):
            for j in range(3):
                for ii in range(3):
                    nb = (bx + ii - 1, by + j - 1, bz + k - 1)
                    if nb in rank:
                        conn[i, ii + 3 * j + 9 * k] = rank[nb]
    return len(order), origins, masks.reshape(-1), conn.reshape(-1)


def random_active(dim, seed, p=0.05):
    rng = np.random.default_rng(seed)
    X, Y, Z = dim
    a = np.zeros((Z, Y, X), np.uint8)
    # sparse random voxels plus a few dense boxes: empty, partial and full blocks
    a[rng.random((Z, Y, X)) < p] = 1
    a[2:12, 3:20, 1:9] = 1
    return a.reshape(-1)


@pytest.mark.parametrize("dim", [(17, 33, 71), (16, 16, 16), (9, 8, 25)])
def test_oracle_builder_brute_force(orc, dim):
    act = random_active(dim, 11)
    n, nact, o, m, c = orc.bgrid_build(dim, act)
    bn, bo, bm, bc = brute_blocks(dim, act)
    assert n == bn and nact == int(act.sum())
    assert np.array_equal(o, bo) and np.array_equal(m, bm) and np.array_equal(c, bc)


def test_neighbour_global_index_property(orc):
    dim = (17, 33, 71)
    X, Y, Z = dim
    act = random_active(dim, 12, p=0.3)
    ob = orc.BlockOracle(dim, act)
    coords = ob.slot_coords()
    active = ob.active_slots()
    a3 = act.reshape(Z, Y, X).astype(bool)
    conn = ob.conn.reshape(-1, 27)
    masks = ob.masks.reshape(-1, 16)
    slots = np.nonzero(active)[0]
    rng = np.random.default_rng(0)
    for s in rng.choice(slots, size=min(4000, slots.size), replace=False):
        b, v = divmod(int(s), 512)
        x, y, z = v & 7, (v >> 3) & 7, v >> 6
        for d in range(27):
            dx, dy, dz = d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1
            nx_, ny_, nz_ = x + dx, y + dy, z + dz
            fx = -1 if nx_ < 0 else (1 if nx_ > 7 else 0)
            fy = -1 if ny_ < 0 else (1 if ny_ > 7 else 0)
            fz = -1 if nz_ < 0 else (1 if nz_ > 7 else 0)
            t = conn[b, (fx + 1) + 3 * (fy + 1) + 9 * (fz + 1)]
            nv = (nx_ & 7) + 8 * (ny_ & 7) + 64 * (nz_ & 7)
            found = t != 0xFFFFFFFF and (masks[t, nv >> 5] >> (nv & 31)) & 1
            g = coords[s] + (dx, dy, dz)
            inside = (g >= 0).all() and g[0] < X and g[1] < Y and g[2] < Z
            expect = inside and a3[g[2], g[1], g[0]]
            assert bool(found) == bool(expect)
            if found:
                assert np.array_equal(coords[t * 512 + nv], g)


def test_library_builder_matches_oracle(orc):
    from neon_amd import abi
    lib = abi.lib
    for dim, per in [((17, 33, 71), 0), ((40, 24, 32), 1), ((16, 16, 16), 7)]:
        act = random_active(dim, 21, p=0.2)
        n, nact, o, m, c = orc.bgrid_build(dim, act, per)
        d = (C.c_int32 * 3)(*dim)
        h = C.c_void_p()
        abi.check(lib.nx_bgrid_create(d, act.ctypes.data, per, C.byref(h)))
        nb, na = C.c_uint32(), C.c_uint64()
        abi.check(lib.nx_bgrid_info(h, C.byref(nb), C.byref(na)))
        assert nb.value == n and na.value == nact
        lo, lm, lc = (np.empty(3 * n, np.int32), np.empty(16 * n, np.uint32),
                      np.empty(27 * n, np.uint32))
        abi.check(lib.nx_bgrid_tables(h, lo.ctypes.data, lm.ctypes.data, lc.ctypes.data))
        lib.nx_bgrid_destroy(h)
        assert np.array_equal(lo, o) and np.array_equal(lm, m) and np.array_equal(lc, c)
        # callback form (the reference's activeCellLambda)
        a3 = act.reshape(dim[2], dim[1], dim[0])
        CB = C.CFUNCTYPE(C.c_int, C.c_int32, C.c_int32, C.c_int32, C.c_void_p)
        cb = CB(lambda x, y, z, u: int(a3[z, y, x]))
        h2 = C.c_void_p()
        abi.check(lib.nx_bgrid_create_cb(d, C.cast(cb, C.c_void_p), None, per, C.byref(h2)))
        lc2 = np.empty(27 * n, np.uint32)
        abi.check(lib.nx_bgrid_tables(h2, None, None, lc2.ctypes.data))
        lib.nx_bgrid_destroy(h2)
        assert np.array_equal(lc2, c)


def test_library_builder_errors(orc):
    from neon_amd import abi
    d = (C.c_int32 * 3)(12, 16, 16)
    act = np.ones(12 * 16 * 16, np.uint8)
    h = C.c_void_p()
    code = abi.lib.nx_bgrid_create(d, act.ctypes.data, 1, C.byref(h))  # periodic x, 12 % 8 != 0
    assert code == abi.NX_ERR_INVALID
    assert b"multiple" in abi.lib.nx_last_error()


def to_slots(ob, dense_arr):
    """Dense [.., Z, Y, X] array -> block slot order (b*512+v)."""
    X, Y, Z = ob.dim
    cz = ob.slot_coords()
    act = ob.active_slots()
    out = np.zeros(dense_arr.shape[:-3] + (cz.shape[0],), dense_arr.dtype)
    xs, ys, zs = (np.clip(cz[:, 0], 0, X - 1), np.clip(cz[:, 1], 0, Y - 1),
                  np.clip(cz[:, 2], 0, Z - 1))
    out[..., act] = dense_arr[..., zs, ys, xs][..., act]
    return out


@pytest.mark.parametrize("storage,compute", [(0, 0), (1, 1), (1, 2), (3, 1)])
def test_block_step_equals_dense_step_cavity(orc, storage, compute):
    N = 20
    act = np.ones(N ** 3, np.uint8)
    dn = orc.DenseOracle((N, N, N), storage, compute)
    dn.setup_cavity(0.05)
    dn.run(15, 1.7)
    ob = orc.BlockOracle((N, N, N), act, storage, compute)
    ob.setup_cavity(0.05)
    ob.run(15, 1.7)
    dense = dn.fin.reshape(19, N, N, N)
    blk = ob.fin.reshape(-1, 19, 512).transpose(1, 0, 2).reshape(19, -1)
    ref = to_slots(ob, dense)
    keep = ob.active_slots()
    assert np.array_equal(blk[:, keep], ref[:, keep])


def test_block_step_equals_dense_step_porous_periodic_forced(orc):
    """Spheres removed from the fluid, periodic x, body force: inactive voxels
    are bounce-back links with a zero wall term in both paths."""
    dim = (32, 24, 24)
    X, Y, Z = dim
    rng = np.random.default_rng(7)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    solid = np.zeros((Z, Y, X), bool)
    for _ in range(6):
        c = rng.random(3) * (X, Y, Z)
        r = rng.uniform(2, 5)
        solid |= (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r
    solid[:, :8, :8] = True    # a whole absent block column
    act = (~solid).astype(np.uint8).reshape(-1)
    g = (2e-5, 0.0, 0.0)
    # dense: solid voxels are 'undefined' cells (never bulk, populations 0)
    dn = orc.DenseOracle(dim, orc.F32, orc.COMPUTE_F32, periodic=1)
    cls = np.where(solid, orc.UNDEFINED, orc.BULK).astype(np.uint32).reshape(-1)
    dn.set_classes(cls)
    ob = orc.BlockOracle(dim, act, orc.F32, orc.COMPUTE_F32, periodic=1)
    ob.set_classes(np.where(ob.active_slots(), orc.BULK, orc.UNDEFINED).astype(np.uint32))
    dn.run(12, 1.4, force=g)
    ob.run(12, 1.4, force=g)
    keep = ob.active_slots()
    blk = ob.fin.reshape(-1, 19, 512).transpose(1, 0, 2).reshape(19, -1)
    ref = to_slots(ob, dn.fin.reshape(19, Z, Y, X))
    assert np.array_equal(blk[:, keep], ref[:, keep])
    # open links exist (solid neighbours, y/z domain faces)
    assert ((ob.flags[keep] & orc.FLAG_OPEN) != 0).sum() > 100
