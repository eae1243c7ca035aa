"""ctypes front end of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / the timed CPU baseline,
never as a product code path.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liblbm_oracle.so"

F64, F32, F32DEV, F16DEV = range(4)
COMPUTE_F64, COMPUTE_F32, COMPUTE_REF = range(3)
BOUNCEBACK, MOVINGWALL, BULK, UNDEFINED = range(4)
CLASS_SHIFT = 24
MASK_BITS = 0x7FDFF
FLAG_OPEN = 1 << 28
FLAG_MIXED = 1 << 29
ZH_SHIFT = 19
ZH_VELOCITY = 8          # added to a face code (bit 22 of the flag word)
NP_STORAGE = {F64: np.float64, F32: np.float32, F32DEV: np.float32, F16DEV: np.uint16}


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class Dense(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("storage", C.c_int32), ("compute", C.c_int32),
                ("periodic", C.c_int32)]


class Block(C.Structure):
    _fields_ = [("dim", C.c_int32 * 3), ("storage", C.c_int32), ("compute", C.c_int32),
                ("periodic", C.c_int32), ("n_blocks", C.c_uint32), ("origins", C.c_void_p),
                ("masks", C.c_void_p), ("conn", C.c_void_p)]


_P = C.c_void_p
_D = C.c_double
_SIG = {
    "or_lattice": (None, [_P, _P, _P]),
    "or_f32_to_f16": (C.c_uint16, [C.c_float]),
    "or_f16_to_f32": (C.c_float, [C.c_uint16]),
    "or_set_threads": (None, [C.c_int]),
    "or_get_threads": (C.c_int, []),
    "or_cavity_flags_dense": (None, [_P, _P]),
    "or_fill_dense": (None, [_P, _P, _P, _D, _P]),
    "or_wall_mask_dense": (None, [_P, _P, _P]),
    "or_step_dense": (None, [_P, _P, _P, _P, _D, _P]),
    "or_step_dense_zouhe": (None, [_P, _P, _P, _P, _D, _P]),
    "or_moments_dense": (None, [_P, _P, _P, _P, _P]),
    "or_mass_dense": (_D, [_P, _P, _P]),
    "or_residual_dense": (None, [_P, _P, _P, _P, _P]),
    "or_run_dense": (None, [_P, _P, _P, _P, _D, _P, C.c_int]),
    "or_bgrid_count": (C.c_uint32, [_P, _P]),
    "or_bgrid_build": (C.c_uint64, [_P, _P, C.c_int32, _P, _P, _P]),
    "or_cavity_flags_block": (None, [_P, _P]),
    "or_fill_block": (None, [_P, _P, _P, _D, _P]),
    "or_wall_mask_block": (None, [_P, _P, _P]),
    "or_step_block": (None, [_P, _P, _P, _P, _D, _P]),
    "or_moments_block": (None, [_P, _P, _P, _P, _P]),
    "or_mass_block": (_D, [_P, _P, _P]),
    "or_run_block": (None, [_P, _P, _P, _P, _D, _P, C.c_int]),
    "or_mres_dims": (None, [_P, C.c_int, _P]),
    "or_mres_build": (C.c_int, [_P, C.c_int, _P, _P]),
    "or_mres_setup": (None, [_P, C.c_int, C.c_int, C.c_int, _D, _P, _P, _P, _P]),
    "or_mres_setup_cells": (None, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
    "or_mres_omega": (_D, [_D, C.c_int, C.c_int]),
    "or_mres_step": (None, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _D]),
    "or_kbc27_collide": (None, [C.c_int, _P, _P, _D]),
}


def _load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    for k, (r, a) in _SIG.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def lattice():
    c = np.zeros((19, 3), np.int32)
    w = np.zeros(19, np.float64)
    o = np.zeros(19, np.int32)
    lib.or_lattice(_ptr(c), _ptr(w), _ptr(o))
    return c, w, o


def set_threads(n: int) -> None:
    lib.or_set_threads(int(n))


class DenseOracle:
    """Oracle state for one dense grid: flags + two population fields."""

    def __init__(self, dim, storage=F64, compute=COMPUTE_F64, periodic=0):
        self.dim = tuple(int(v) for v in dim)
        self.d = Dense()
        for a in range(3):
            self.d.dim[a] = self.dim[a]
        self.d.storage, self.d.compute, self.d.periodic = storage, compute, periodic
        self.n = self.dim[0] * self.dim[1] * self.dim[2]
        self.flags = np.zeros(self.n, np.uint32)
        self.pop = [np.zeros(19 * self.n, NP_STORAGE[storage]) for _ in range(2)]
        self.parity = 0

    @property
    def storage(self):
        return self.d.storage

    def setup_cavity(self, ulb, ulid=(1.0, 0.0, 0.0)):
        lib.or_cavity_flags_dense(C.byref(self.d), _ptr(self.flags))
        self._fill_and_mask(ulb, ulid)

    def set_classes(self, cls, ulb=0.0, ulid=(1.0, 0.0, 0.0), zouhe=None):
        """zouhe: optional per-cell Zou-He code (face 1..6, + 8 for a
        velocity face), stored in bits 19..22 of bulk cells' flag words."""
        self.flags[:] = np.asarray(cls, np.uint32).reshape(-1) << CLASS_SHIFT
        if zouhe is not None:
            self.flags[:] |= (np.asarray(zouhe, np.uint32).reshape(-1) & 15) << ZH_SHIFT
        self._fill_and_mask(ulb, ulid)

    def _fill_and_mask(self, ulb, ulid):
        for p in self.pop:
            lib.or_fill_dense(C.byref(self.d), _ptr(p), _ptr(self.flags), float(ulb), _d3(ulid))
        lib.or_wall_mask_dense(C.byref(self.d), _ptr(self.flags), _ptr(self.flags))
        self.parity = 0

    @property
    def fin(self):
        return self.pop[self.parity]

    def run(self, nsteps, omega, force=(0.0, 0.0, 0.0)):
        for _ in range(int(nsteps)):
            lib.or_step_dense(C.byref(self.d), _ptr(self.pop[self.parity]),
                              _ptr(self.pop[self.parity ^ 1]), _ptr(self.flags), float(omega),
                              _d3(force))
            self.parity ^= 1

    def run_zouhe(self, nsteps, omega, values):
        """Steps with Zou-He cells; values[face - 1] = prescribed density or
        inward normal velocity (LbmContainers.h:145-287 of lbm-flow-over-sphere)."""
        v = (C.c_double * 6)(*[float(x) for x in values])
        for _ in range(int(nsteps)):
            lib.or_step_dense_zouhe(C.byref(self.d), _ptr(self.pop[self.parity]),
                                    _ptr(self.pop[self.parity ^ 1]), _ptr(self.flags), float(omega), v)
            self.parity ^= 1

    def moments(self):
        mt = np.float64 if self.d.storage == F64 else np.float32
        rho = np.zeros(self.n, mt)
        u = np.zeros(3 * self.n, mt)
        lib.or_moments_dense(C.byref(self.d), _ptr(self.fin), _ptr(self.flags), _ptr(rho), _ptr(u))
        X, Y, Z = self.dim
        return rho.reshape(Z, Y, X), u.reshape(3, Z, Y, X)

    def mass(self):
        return float(lib.or_mass_dense(C.byref(self.d), _ptr(self.fin), _ptr(self.flags)))

    def residual(self, u_prev: np.ndarray):
        out = np.zeros(2, np.float64)
        lib.or_residual_dense(C.byref(self.d), _ptr(self.fin), _ptr(self.flags), _ptr(u_prev),
                              _ptr(out))
        return float(out[0]), float(out[1])

    def decoded(self) -> np.ndarray:
        """Populations f (not the stored encoding) as float64 [19, n]."""
        return decode(self.fin, self.d.storage, self.flags).reshape(19, self.n)


def decode(raw: np.ndarray, storage: int, flags: np.ndarray) -> np.ndarray:
    """Stored encoding -> f (float64) for every slot, [19, n] flattened."""
    n = flags.size
    if storage == F16DEV:
        v = np.asarray(raw).view(np.float16).astype(np.float32).reshape(19, n)
    else:
        v = np.asarray(raw).reshape(19, n)
    if storage in (F32DEV, F16DEV):
        _, w, _ = lattice()
        bulk = ((flags >> CLASS_SHIFT) & 3) == BULK
        wf = w.astype(np.float32)[:, None]
        v = np.where(bulk[None, :], (wf + v.astype(np.float32)), v)
    return v.astype(np.float64).reshape(-1)


def bgrid_build(dim, active: np.ndarray, periodic=0):
    d = (C.c_int32 * 3)(*[int(v) for v in dim])
    act = np.ascontiguousarray(active, np.uint8).reshape(-1)
    n = int(lib.or_bgrid_count(d, _ptr(act)))
    origins = np.zeros(3 * max(n, 1), np.int32)
    masks = np.zeros(16 * max(n, 1), np.uint32)
    conn = np.zeros(27 * max(n, 1), np.uint32)
    nact = int(lib.or_bgrid_build(d, _ptr(act), int(periodic), _ptr(origins), _ptr(masks),
                                  _ptr(conn)))
    return n, nact, origins[:3 * n], masks[:16 * n], conn[:27 * n]


class BlockOracle:
    def __init__(self, dim, active, storage=F64, compute=COMPUTE_F64, periodic=0):
        self.dim = tuple(int(v) for v in dim)
        (self.n_blocks, self.n_active, self.origins, self.masks,
         self.conn) = bgrid_build(dim, active, periodic)
        self.d = Block()
        for a in range(3):
            self.d.dim[a] = self.dim[a]
        self.d.storage, self.d.compute, self.d.periodic = storage, compute, periodic
        self.d.n_blocks = self.n_blocks
        self.d.origins, self.d.masks, self.d.conn = (_ptr(self.origins), _ptr(self.masks),
                                                     _ptr(self.conn))
        self.n = self.n_blocks * 512
        self.flags = np.zeros(self.n, np.uint32)
        self.pop = [np.zeros(19 * self.n, NP_STORAGE[storage]) for _ in range(2)]
        self.parity = 0

    def setup_cavity(self, ulb, ulid=(1.0, 0.0, 0.0)):
        lib.or_cavity_flags_block(C.byref(self.d), _ptr(self.flags))
        self._fill_and_mask(ulb, ulid)

    def set_classes(self, cls_slots, ulb=0.0, ulid=(1.0, 0.0, 0.0)):
        self.flags[:] = np.asarray(cls_slots, np.uint32).reshape(-1) << CLASS_SHIFT
        self._fill_and_mask(ulb, ulid)

    def _fill_and_mask(self, ulb, ulid):
        for p in self.pop:
            lib.or_fill_block(C.byref(self.d), _ptr(p), _ptr(self.flags), float(ulb), _d3(ulid))
        lib.or_wall_mask_block(C.byref(self.d), _ptr(self.flags), _ptr(self.flags))
        self.parity = 0

    @property
    def fin(self):
        return self.pop[self.parity]

    def run(self, nsteps, omega, force=(0.0, 0.0, 0.0)):
        for _ in range(int(nsteps)):
            lib.or_step_block(C.byref(self.d), _ptr(self.pop[self.parity]),
                              _ptr(self.pop[self.parity ^ 1]), _ptr(self.flags), float(omega),
                              _d3(force))
            self.parity ^= 1

    def moments(self):
        mt = np.float64 if self.d.storage == F64 else np.float32
        rho = np.zeros(self.n, mt)
        u = np.zeros(3 * self.n, mt)
        lib.or_moments_block(C.byref(self.d), _ptr(self.fin), _ptr(self.flags), _ptr(rho), _ptr(u))
        return rho, u

    def mass(self):
        return float(lib.or_mass_block(C.byref(self.d), _ptr(self.fin), _ptr(self.flags)))

    def slot_coords(self):
        """Global (x,y,z) of every voxel slot b*512+v, [n, 3]."""
        v = np.arange(512)
        loc = np.stack([v & 7, (v >> 3) & 7, v >> 6], axis=1)
        o = self.origins.reshape(-1, 3)
        return (o[:, None, :] + loc[None, :, :]).reshape(-1, 3)

    def active_slots(self):
        bits = np.unpackbits(self.masks.view(np.uint8), bitorder="little")
        return bits.reshape(-1).astype(bool)


MR_ACTIVE, MR_REFINED, MR_CLS_SHIFT = 1, 2, 2
BGK19, KBC27 = 0, 1  # lattice and collision of the refinement oracle

# D3Q27 of apps/lbmMultiRes/lattice.h:14-75 (z fastest, then y, then x, each
# over 0, -1, +1)
D3Q27_C = np.array([(x, y, z) for x in (0, -1, 1) for y in (0, -1, 1) for z in (0, -1, 1)], np.int64)


def kbc27_collide(f, omega, storage=F64):
    """The oracle's KBC D3Q27 collision of single voxels: f [..., 27]."""
    dt = NP_STORAGE[storage]
    f = np.ascontiguousarray(f, dt)
    out = np.empty_like(f)
    flat_in, flat_out = f.reshape(-1, 27), out.reshape(-1, 27)
    for k in range(flat_in.shape[0]):
        lib.or_kbc27_collide(storage, _ptr(flat_in[k]), _ptr(flat_out[k]), float(omega))
    return out


class MresOracle:
    """Grid refinement LBM over `depth` levels (mres_oracle.c, restating
    apps/lbmMultiRes, BGK D3Q19).  MresOracle(dim, [lam_0, .., lam_{D-1}],
    storage) or, for two levels, MresOracle(dim, lam0, lam1, storage): lam_l
    is the activation lambda of level l at every level-l voxel (evaluated at
    its level-0 origin v 2^l), uint8, x fastest.  Fields are dense per level,
    [19][z][y][x]; st[l] the per-voxel state bytes (st0 / st1, dim1 name
    levels 0 and 1)."""

    def __init__(self, dim, lams, *rest, lattice=BGK19):
        if isinstance(lams, (list, tuple)):
            storage = rest[0] if rest else F64
        else:  # (dim, lam0, lam1[, storage])
            lams = [lams, rest[0]]
            storage = rest[1] if len(rest) > 1 else F64
        self.lattice = lattice
        self.Q = 27 if lattice == KBC27 else 19
        self.depth = len(lams)
        assert 2 <= self.depth <= 5
        self.dim = tuple(int(v) for v in dim)
        self._d0 = (C.c_int32 * 3)(*self.dim)
        dims = (C.c_int32 * (3 * self.depth))()
        lib.or_mres_dims(self._d0, self.depth, dims)
        self.dims = [tuple(int(v) for v in dims[3 * l:3 * l + 3]) for l in range(self.depth)]
        self.dim1 = self.dims[1]
        self.n = [int(np.prod(d)) for d in self.dims]
        self.n0, self.n1 = self.n[0], self.n[1]
        self.storage = storage
        lams = [np.ascontiguousarray(a, np.uint8).reshape(-1) for a in lams]
        assert all(a.size == n for a, n in zip(lams, self.n))
        self.st = [np.zeros(n, np.uint8) for n in self.n]
        self._lams = lams
        self.balanced = int(lib.or_mres_build(self._d0, self.depth, self._ptrs(lams), self._ptrs(self.st)))

    @staticmethod
    def _ptrs(arrays):
        return (C.c_void_p * len(arrays))(*[None if a is None else a.ctypes.data for a in arrays])

    @property
    def st0(self):
        return self.st[0]

    @property
    def st1(self):
        return self.st[1]

    def active(self, level):
        return (self.st[level] & MR_ACTIVE) != 0

    def refined(self, level=1):
        return (self.st[level] & MR_REFINED) != 0

    def setup_cavity(self, ulb):
        dt = NP_STORAGE[self.storage]
        self.fin = [np.zeros(self.Q * n, dt) for n in self.n]
        self.fout = [np.zeros(self.Q * n, dt) for n in self.n]
        self.ss = [None] + [np.zeros(self.Q * n, np.float32) for n in self.n[1:]]
        self.ss1 = self.ss[1]
        lib.or_mres_setup(self._d0, self.depth, self.lattice, self.storage, float(ulb), self._ptrs(self.st),
                          self._ptrs(self.fin), self._ptrs(self.fout), self._ptrs(self.ss))

    def setup_cells(self, cls, uw):
        """Set-up with the class of every leaf per level (cls[l], uint8 per
        level-l voxel) and the moving-wall velocity uw (lid or inlet)."""
        dt = NP_STORAGE[self.storage]
        self.fin = [np.zeros(self.Q * n, dt) for n in self.n]
        self.fout = [np.zeros(self.Q * n, dt) for n in self.n]
        self.ss = [None] + [np.zeros(self.Q * n, np.float32) for n in self.n[1:]]
        self.ss1 = self.ss[1]
        cls = [np.ascontiguousarray(c, np.uint8).reshape(-1) for c in cls]
        assert all(c.size == n for c, n in zip(cls, self.n))
        lib.or_mres_setup_cells(self._d0, self.depth, self.lattice, self.storage, self._ptrs(cls),
                                (C.c_double * 3)(*[float(v) for v in uw]), self._ptrs(self.st),
                                self._ptrs(self.fin), self._ptrs(self.fout), self._ptrs(self.ss))

    def classes(self, level):
        return (self.st[level] >> MR_CLS_SHIFT) & 3

    def run(self, n, omega0):
        st, fi, fo, ss = self._ptrs(self.st), self._ptrs(self.fin), self._ptrs(self.fout), self._ptrs(self.ss)
        for _ in range(int(n)):
            lib.or_mres_step(self._d0, self.depth, self.lattice, self.storage, st, fi, fo, ss, float(omega0))

    @staticmethod
    def omega(omega0, level, num_levels=2):
        return float(lib.or_mres_omega(float(omega0), int(level), int(num_levels)))
