"""ctypes binding of the C ABI declared in include/neon_amd.h.

The shared library neon_amd/lib/libneon_amd.so must have been built
(``python neon_amd/build.py``); there is no fallback: importing this module
without it raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libneon_amd.so"

NX_OK, NX_ERR_INVALID, NX_ERR_HIP, NX_ERR_OOM, NX_ERR_UNSUPPORTED, NX_ERR_NOT_INIT = range(6)
NX_F64, NX_F32, NX_F32DEV, NX_F16DEV = range(4)
NX_COMPUTE_F64, NX_COMPUTE_F32, NX_COMPUTE_REF = range(3)
NX_CELL_BOUNCEBACK, NX_CELL_MOVINGWALL, NX_CELL_BULK, NX_CELL_UNDEFINED = range(4)
NX_CLASS_SHIFT = 24
NX_MASK_BITS = 0x7FDFF
NX_FLAG_OPEN = 1 << 28
NX_FLAG_MIXED = 1 << 29
NX_ZH_SHIFT = 19
NX_ZH_XMIN, NX_ZH_XMAX, NX_ZH_YMIN, NX_ZH_YMAX, NX_ZH_ZMIN, NX_ZH_ZMAX = range(1, 7)
NX_ZH_VELOCITY_CODE = 8   # added to a face code: bit 22 of the flag word
NX_PERIODIC_X, NX_PERIODIC_Y, NX_PERIODIC_Z = 1, 2, 4
NX_CONV_RELU, NX_CONV_OUT_F32 = 1, 2

STORAGE_BYTES = {NX_F64: 8, NX_F32: 4, NX_F32DEV: 4, NX_F16DEV: 2}
STORAGE_NAMES = {"f64": NX_F64, "f32": NX_F32, "f32dev": NX_F32DEV, "f16dev": NX_F16DEV}
COMPUTE_NAMES = {"f64": NX_COMPUTE_F64, "f32": NX_COMPUTE_F32, "ref": NX_COMPUTE_REF}


class NeonError(RuntimeError):
    """Error returned by the C ABI (the counterpart of Neon::NeonException)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[nx error {code}] {msg}")
        self.code = code


class DenseDesc(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("storage", C.c_int32), ("compute", C.c_int32),
                ("periodic", C.c_int32)]


class BlockDesc(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("storage", C.c_int32), ("compute", C.c_int32),
                ("periodic", C.c_int32), ("n_blocks", C.c_uint32), ("reserved", C.c_uint32),
                ("origins", C.c_void_p), ("masks", C.c_void_p), ("conn", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("wg", C.c_int32), ("cpt", C.c_int32), ("load_policy", C.c_int32),
                ("xcd_remap", C.c_int32), ("vx", C.c_int32), ("boundary", C.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class LbmParams(C.Structure):
    _fields_ = [("omega", C.c_double), ("force", C.c_double * 3)]


class UNetDesc(C.Structure):
    _fields_ = [("levels", C.c_int32), ("width", C.c_int32 * 4), ("size", C.c_int32),
                ("max_batch", C.c_int32), ("in_ch", C.c_int32), ("out_ch", C.c_int32),
                ("reserved", C.c_int32), ("seed", C.c_uint64)]


_P = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t
_D = C.c_double
_DP = C.POINTER(DenseDesc)
_BP = C.POINTER(BlockDesc)
_PP = C.POINTER(LbmParams)
_U32 = C.c_uint32
_I32 = C.c_int32

# name -> (restype, argtypes); every entry point of include/neon_amd.h
SIGNATURES = {
    "nx_init": (_I, [_I]),
    "nx_last_error": (C.c_char_p, []),
    "nx_device_count": (_I, [C.POINTER(_I)]),
    "nx_device_name": (_I, [C.c_char_p, _I]),
    "nx_sync": (_I, []),
    "nx_stream_create": (_I, [C.POINTER(_P)]),
    "nx_stream_sync": (_I, [_P]),
    "nx_stream_destroy": (_I, [_P]),
    "nx_event_create": (_I, [C.POINTER(_P)]),
    "nx_event_record": (_I, [_P, _P]),
    "nx_event_sync": (_I, [_P]),
    "nx_event_elapsed_ms": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "nx_event_destroy": (_I, [_P]),
    "nx_alloc": (_I, [_SZ, C.POINTER(_P)]),
    "nx_free": (_I, [_P]),
    "nx_memset": (_I, [_P, _I, _SZ, _P]),
    "nx_h2d": (_I, [_P, _P, _SZ, _P]),
    "nx_d2h": (_I, [_P, _P, _SZ, _P]),
    "nx_d2d": (_I, [_P, _P, _SZ, _P]),
    "nx_host_alloc": (_I, [_SZ, C.POINTER(_P)]),
    "nx_host_free": (_I, [_P]),
    "nx_copy_kernel": (_I, [_P, _P, _SZ, _P]),
    "nx_shifted_copy": (_I, [_P, _P, _P, _I32, _P]),
    "nx_d3q19_get_tuning_dense": (_I, [_DP, C.POINTER(Tuning)]),
    "nx_d3q19_set_tuning_dense": (_I, [_DP, C.POINTER(Tuning)]),
    "nx_d3q19_iteration_dense": (_I, [_DP, _P, _P, _P, _PP, C.POINTER(_P)]),
    "nx_container_run": (_I, [_P, _P]),
    "nx_container_run_pair": (_I, [_P, _P, _I32, _P]),
    "nx_container_info": (_I, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                               C.POINTER(C.c_int)]),
    "nx_container_destroy": (_I, [_P]),
    "nx_container_autotune": (_I, [_P, _P, C.POINTER(Tuning), C.POINTER(C.c_float)]),
    "nx_pop_bytes_dense": (_SZ, [_DP]),
    "nx_pop_bytes_block": (_SZ, [_BP]),
    "nx_cavity_flags_dense": (_I, [_DP, _P, _P]),
    "nx_d3q19_fill_dense": (_I, [_DP, _P, _P, _D, C.POINTER(_D), _P]),
    "nx_d3q19_wall_mask_dense": (_I, [_DP, _P, _P, _P]),
    "nx_d3q19_step_dense": (_I, [_DP, _P, _P, _P, _PP, _P]),
    "nx_d3q19_step_dense_zouhe": (_I, [_DP, _P, _P, _P, _PP, C.POINTER(C.c_double), _P]),
    "nx_d3q19_moments_dense": (_I, [_DP, _P, _P, _P, _P, _P]),
    "nx_d3q19_mass_dense": (_I, [_DP, _P, _P, _P, _P]),
    "nx_d3q19_residual_dense": (_I, [_DP, _P, _P, _P, _P, _P]),
    "nx_d3q19_residual_dense_noalloc": (_I, [_DP, _P, _P, _P, _P, _P]),
    "nx_bgrid_create": (_I, [C.POINTER(_I32), _P, _I32, C.POINTER(_P)]),
    "nx_bgrid_create_cb": (_I, [C.POINTER(_I32), _P, _P, _I32, C.POINTER(_P)]),
    "nx_bgrid_info": (_I, [_P, C.POINTER(_U32), C.POINTER(C.c_uint64)]),
    "nx_bgrid_tables": (_I, [_P, _P, _P, _P]),
    "nx_bgrid_destroy": (_I, [_P]),
    "nx_cavity_flags_block": (_I, [_BP, _P, _P]),
    "nx_d3q19_fill_block": (_I, [_BP, _P, _P, _D, C.POINTER(_D), _P]),
    "nx_d3q19_wall_mask_block": (_I, [_BP, _P, _P, _P]),
    "nx_d3q19_step_block": (_I, [_BP, _P, _P, _P, _PP, _P]),
    "nx_block_tiles_create": (_I, [_BP, C.POINTER(_P)]),
    "nx_block_tiles_info": (_I, [_P, C.POINTER(_U32)]),
    "nx_block_tiles_destroy": (_I, [_P]),
    "nx_d3q19_step_block_tiled": (_I, [_BP, _P, _P, _P, _P, _PP, _P]),
    "nx_d3q19_set_block_remap": (_I, [_I]),
    "nx_d3q19_set_block_pairs": (_I, [_I]),
    "nx_geometry_spheres": (_I, [C.POINTER(_I32), C.c_uint32, _D, _D, C.c_uint64, _I32, _P]),
    "nx_geometry_cavity": (_I, [C.POINTER(_I32), C.c_uint32, _I32, _I32, C.c_uint64, _P]),
    "nx_batch_create": (_I, [C.c_uint32, _P, _P, _P, _P, _P, _I32, C.POINTER(_P)]),
    "nx_batch_info": (_I, [_P, C.POINTER(_U32), C.POINTER(C.c_uint64), C.POINTER(_U32)]),
    "nx_batch_member": (_I, [_P, _U32, _I, _DP, C.POINTER(_P), C.POINTER(_P)]),
    "nx_batch_init": (_I, [_P, _P, _P, _P]),
    "nx_batch_step": (_I, [_P, _I32, _P]),
    "nx_batch_run": (_I, [_P, _I32, _I32, _D, _P, _P, C.POINTER(_U32)]),
    "nx_batch_set_launch": (_I, [_P, _I32, _I32]),
    "nx_batch_set_variant": (_I, [_P, _I32, _I32]),
    "nx_batch_get_launch": (_I, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)]),
    "nx_batch_autotune": (_I, [_P, _P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(C.c_float)]),
    "nx_batch_destroy": (_I, [_P]),
    "nx_mgrid_create_levels": (_I, [_P, _I32, _P, _I32, _I32, C.POINTER(_P)]),
    "nx_mgrid_populations": (_I, [_P, _P]),
    "nx_mgrid_create": (_I, [_P, _P, _P, _I32, C.POINTER(_P)]),
    "nx_mgrid_depth": (_I, [_P, _P, _P]),
    "nx_mgrid_state_level": (_I, [_P, _I32, _P]),
    "nx_mgrid_info": (_I, [_P, _P, _P, _P, _P, _P]),
    "nx_mgrid_state": (_I, [_P, _P, _P]),
    "nx_mgrid_cavity": (_I, [_P, _D]),
    "nx_mgrid_setup_cells": (_I, [_P, _P, _P]),
    "nx_mgrid_step": (_I, [_P, _D, _I32, _P]),
    "nx_mgrid_fields": (_I, [_P, _I32, _I32, _P]),
    "nx_mgrid_destroy": (_I, [_P]),
    "nx_batch_grid_size": (_I, [_U32, _I32, _I32, C.POINTER(_U32)]),
    "nx_batch_launch_grid": (_I, [_P, C.POINTER(_U32), C.POINTER(_I32)]),
    "nx_batch_constrain": (_I, [_P, _P, _I32, _P, _P]),
    "nx_conv3d_fwd_bf16": (_I, [_P, _P, _P, _P, C.POINTER(_I32), _I32, _I32, _I32, _I32, _P]),
    "nx_conv3d_bwd_data_bf16": (_I, [_P, _P, _P, C.POINTER(_I32), _I32, _I32, _I32, _I32, _P]),
    "nx_conv3d_bwd_weight_bf16": (_I, [_P, _P, _P, _P, C.POINTER(_I32), _I32, _I32, _I32, _P]),
    "nx_unet_create": (_I, [C.POINTER(UNetDesc), C.POINTER(_P)]),
    "nx_unet_info": (_I, [_P, C.POINTER(C.c_uint64), C.POINTER(_U32)]),
    "nx_unet_params": (_I, [_P, _P, _I32]),
    "nx_unet_grads": (_I, [_P, _P]),
    "nx_unet_forward": (_I, [_P, _P, _I32, _P, _P]),
    "nx_unet_backward": (_I, [_P, _P, _I32, _P]),
    "nx_unet_train_step": (_I, [_P, _P, _P, _P, _I32, C.c_float, C.POINTER(_D), _P]),
    "nx_unet_destroy": (_I, [_P]),
    "nx_d3q19_moments_block": (_I, [_BP, _P, _P, _P, _P, _P]),
    "nx_d3q19_mass_block": (_I, [_BP, _P, _P, _P, _P]),
    "nx_d3q19_residual_block": (_I, [_BP, _P, _P, _P, _P, _P]),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python neon_amd/build.py` (HIP library for gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(code: int) -> None:
    """Raise NeonError for a non-zero status of the C ABI."""
    if code != NX_OK:
        msg = lib.nx_last_error()
        raise NeonError(code, msg.decode() if msg else "")
