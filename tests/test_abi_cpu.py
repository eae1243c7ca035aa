"""The C ABI library loads on a host without a GPU and exports every entry
point that include/neon_amd.h declares; argument checking happens before any
device work."""
import ctypes as C
import pathlib
import re

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "neon_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nx_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    from neon_amd import abi
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(abi.lib, n)]
    assert not missing, missing
    # and the binding declares a signature for each of them
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)


def test_argument_errors_without_device_work():
    from neon_amd import abi
    d = abi.DenseDesc()
    d.dim[0], d.dim[1], d.dim[2] = 8, 8, 8
    d.storage, d.compute = abi.NX_F16DEV, abi.NX_COMPUTE_F64
    p = abi.LbmParams()
    code = abi.lib.nx_d3q19_step_dense(C.byref(d), C.c_void_p(16), C.c_void_p(32),
                                       C.c_void_p(64), C.byref(p), None)
    assert code == abi.NX_ERR_UNSUPPORTED
    assert b"deviation" in abi.lib.nx_last_error()
    d.storage, d.compute = abi.NX_F32, abi.NX_COMPUTE_F32
    code = abi.lib.nx_d3q19_step_dense(C.byref(d), C.c_void_p(16), C.c_void_p(16),
                                       C.c_void_p(64), C.byref(p), None)
    assert code == abi.NX_ERR_INVALID and b"differ" in abi.lib.nx_last_error()
    d.dim[0] = 0
    assert abi.lib.nx_d3q19_wall_mask_dense(C.byref(d), C.c_void_p(4), C.c_void_p(4),
                                            None) == abi.NX_ERR_INVALID
    b = abi.BlockDesc()
    assert abi.lib.nx_d3q19_step_block(C.byref(b), None, None, None, C.byref(p),
                                       None) == abi.NX_ERR_INVALID
    # Zou-He step: deviation storage and a body force are refused up front
    d.dim[0] = 8
    vals = (C.c_double * 6)(1.0, 1.0, 0, 0, 0, 0)
    d.storage, d.compute = abi.NX_F32DEV, abi.NX_COMPUTE_F32
    assert abi.lib.nx_d3q19_step_dense_zouhe(C.byref(d), C.c_void_p(16), C.c_void_p(32), C.c_void_p(64),
                                             C.byref(p), vals, None) == abi.NX_ERR_UNSUPPORTED
    d.storage = abi.NX_F32
    pf = abi.LbmParams()
    pf.force[0] = 1e-5
    assert abi.lib.nx_d3q19_step_dense_zouhe(C.byref(d), C.c_void_p(16), C.c_void_p(32), C.c_void_p(64),
                                             C.byref(pf), vals, None) == abi.NX_ERR_UNSUPPORTED
    assert abi.lib.nx_d3q19_step_dense_zouhe(C.byref(d), C.c_void_p(16), C.c_void_p(32), C.c_void_p(64),
                                             C.byref(p), None, None) == abi.NX_ERR_INVALID
    # the allocation-free residual checks its arguments like the residual
    assert abi.lib.nx_d3q19_residual_dense_noalloc(C.byref(d), C.c_void_p(16), C.c_void_p(32),
                                                   None, C.c_void_p(64), None) == abi.NX_ERR_INVALID
    # container pair / batch variant: null handles and bad arguments
    assert abi.lib.nx_container_run_pair(None, None, 4, None) == abi.NX_ERR_INVALID
    assert abi.lib.nx_batch_set_variant(None, 1, 1) == abi.NX_ERR_INVALID
    try:
        abi.check(abi.NX_ERR_INVALID)
    except abi.NeonError as e:
        assert e.code == abi.NX_ERR_INVALID
    else:
        raise AssertionError("check() did not raise")


def test_pop_bytes():
    from neon_amd import abi
    d = abi.DenseDesc()
    d.dim[0], d.dim[1], d.dim[2] = 4, 5, 6
    for st, sz in abi.STORAGE_BYTES.items():
        d.storage = st
        assert abi.lib.nx_pop_bytes_dense(C.byref(d)) == 4 * 5 * 6 * 19 * sz
    b = abi.BlockDesc()
    b.n_blocks, b.storage = 3, abi.NX_F16DEV
    assert abi.lib.nx_pop_bytes_block(C.byref(b)) == 3 * 19 * 512 * 2


def test_batch_grid_size():
    """Persistent grid of the batched step: min(tiles, CUs x workgroups per
    CU) in 64-bit arithmetic; 0 workgroups per CU means one per tile."""
    from neon_amd import abi

    def grid(tiles, cus, per_cu):
        g = C.c_uint32()
        assert abi.lib.nx_batch_grid_size(tiles, cus, per_cu, C.byref(g)) == abi.NX_OK
        return g.value

    tiles = 688_128                         # 512 cavities of 32..96 cells, 256-cell tiles
    assert grid(tiles, 256, 0) == tiles     # one workgroup per tile
    assert grid(tiles, 304, 0) == tiles
    assert grid(tiles, 256, 2) == 512
    assert grid(tiles, 256, 16) == 4096
    assert grid(1000, 256, 1 << 30) == 1000   # 256 * 2^30 does not wrap to 0
    assert grid(tiles, 2 ** 31 - 1, 2 ** 31 - 1) == tiles
    assert grid(5, 304, 16) == 5
    assert grid(0, 256, 0) == 1
    g = C.c_uint32()
    assert abi.lib.nx_batch_grid_size(10, 256, -1, C.byref(g)) == abi.NX_ERR_INVALID
    assert abi.lib.nx_batch_grid_size(10, 256, 2, None) == abi.NX_ERR_INVALID
    assert abi.lib.nx_batch_launch_grid(None, C.byref(g), None) == abi.NX_ERR_INVALID
