"""Two-level grid refinement (Neon's mGrid with depth 2) driven through the
C ABI (nx_mgrid_*): the level activation of the reference's lid-driven cavity
(apps/lbmMultiRes/lidDrivenCavity.h:88-271) and the solver handle."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib


def sdf_cube(ids: np.ndarray, dim) -> np.ndarray:
    """Signed distance to the faces of the box [0, dim] mapped to [-1, 1]^3
    (apps/lbmMultiRes/util.h:65-81) for integer voxel positions ids [..., 3]
    (x, y, z), in binary32 like the reference: 0 on the faces, -1 at the
    centre."""
    half = np.asarray(dim, np.float32) * np.float32(0.5)
    p = (ids.astype(np.float32) - half) / half
    d = np.abs(p) - np.float32(1.0)
    dm = np.maximum(d, np.float32(0.0))
    length = np.sqrt((dm * dm).sum(-1, dtype=np.float32)).astype(np.float32)
    return (np.minimum(d.max(-1), np.float32(0.0)) + length).astype(np.float32)


def cavity_lambdas(dim, level_sdf=(0.0, -1.0 / 3.0, -1.0)):
    """Activation of the two levels of the refined cavity: level l holds the
    voxels with level_sdf[l+1] < sdfCube <= level_sdf[l] (lidDrivenCavity.h:
    229-236; sdfCube of the level-0 position against dim - 1).  Returns
    (lam0 at every level-0 voxel, lam1 at every level-1 voxel evaluated at
    its level-0 origin 2v), uint8, x fastest."""
    X, Y, Z = (int(v) for v in dim)
    box = np.array([X - 1, Y - 1, Z - 1])
    lv = [np.float32(v) for v in level_sdf]

    def lam(shape, scale, lo, hi):
        z, y, x = np.meshgrid(*(np.arange(n) for n in shape[::-1]), indexing="ij")
        ids = np.stack([x, y, z], -1) * scale
        s = sdf_cube(ids, box)
        return ((s <= lo) & (s > hi)).astype(np.uint8).reshape(-1)

    d1 = ((X + 1) // 2, (Y + 1) // 2, (Z + 1) // 2)
    return lam((X, Y, Z), 1, lv[0], lv[1]), lam(d1, 2, lv[1], lv[2])


def cavity_parameters(dim, re: float, storage: str = "f64", ulb: float = 0.04):
    """(ulb, omega0) of the refined cavity in the field type T
    (lidDrivenCavity.h:246-251): ulb = 0.04 in T, the characteristic length is
    the top level's dimension (its groups times 2), nu = ulb * length / Re in
    T, omega0 = 1 / (3 nu + 1/2) evaluated in binary64 and stored in T."""
    T = np.float64 if storage == "f64" else np.float32
    u = T(ulb)
    length = T(2 * ((int(dim[0]) + 3) // 4))
    nu = T(u * length / T(re))
    omega = T(1.0 / (3.0 * float(nu) + 0.5))
    return float(u), float(omega)


class MGrid:
    """Two-level refined grid and its LBM fields on the device (nx_mgrid_*).
    lam0 / lam1: activation of level 0 (every level-0 voxel) and level 1
    (every level-1 voxel at its level-0 origin), uint8, x fastest."""

    def __init__(self, dim, lam0, lam1, storage="f64"):
        from .lattice import init, _storage
        self.dim = tuple(int(v) for v in dim)
        self.dim1 = tuple((v + 1) // 2 for v in self.dim)
        self.storage = _storage(storage)
        lam0 = np.ascontiguousarray(lam0, np.uint8).reshape(-1)
        lam1 = np.ascontiguousarray(lam1, np.uint8).reshape(-1)
        assert lam0.size == int(np.prod(self.dim)) and lam1.size == int(np.prod(self.dim1))
        h = C.c_void_p()
        check(lib.nx_mgrid_create((C.c_int32 * 3)(*self.dim), lam0.ctypes.data, lam1.ctypes.data,
                                  self.storage, C.byref(h)))
        self.h = h.value
        self._init = init

    def __del__(self):
        try:
            lib.nx_mgrid_destroy(self.h)
        except Exception:
            pass

    def state(self):
        """(state bytes of level 0, of level 1): bit 0 active, bit 1 refined,
        bits 2..3 the cell class."""
        s0 = np.zeros(int(np.prod(self.dim)), np.uint8)
        s1 = np.zeros(int(np.prod(self.dim1)), np.uint8)
        check(lib.nx_mgrid_state(self.h, s0.ctypes.data, s1.ctypes.data))
        return s0, s1

    def info(self) -> dict:
        nb = (C.c_uint32 * 2)()
        na, nl = (C.c_uint64 * 2)(), (C.c_uint64 * 2)()
        ns, fused = C.c_uint64(), C.c_int32()
        check(lib.nx_mgrid_info(self.h, nb, na, nl, C.byref(ns), C.byref(fused)))
        return {"blocks": list(nb), "active": list(na), "leaf": list(nl), "store_links": int(ns.value),
                "fine_fused": bool(fused.value)}

    def setup_cavity(self, ulb: float) -> None:
        self._init()
        check(lib.nx_mgrid_cavity(self.h, float(ulb)))

    def run(self, n: int, omega0: float, stream=None) -> None:
        check(lib.nx_mgrid_step(self.h, float(omega0), int(n), stream))

    def fields(self, level: int, which: int = 0) -> np.ndarray:
        """fin (which 0) or fout (which 1) of one level, dense [19 * n]."""
        n = int(np.prod(self.dim1 if level else self.dim))
        dt = np.float64 if self.storage == abi.NX_F64 else np.float32
        out = np.zeros(19 * n, dt)
        check(lib.nx_mgrid_fields(self.h, int(level), int(which), out.ctypes.data))
        return out
