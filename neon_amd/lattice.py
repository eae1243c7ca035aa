"""Python-side plumbing over the C ABI for tests and the benchmark.

Mirrors the objects of the reference's cavity driver
(benchmarks/lbm-lid-driven-cavity-flow/src/RunCavityTwoPop.cu:17-276 and
LbmIteration.h:17-101): two population fields that ping-pong, a cell-type
field with the wall-neighbour mask, and moments / mass / residual read-backs.
All compute goes through libneon_amd.so; nothing here evaluates the lattice
Boltzmann step on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib

_NP_STORAGE = {abi.NX_F64: np.float64, abi.NX_F32: np.float32, abi.NX_F32DEV: np.float32,
               abi.NX_F16DEV: np.float16}

_initialised: set[int] = set()


def init(device: int = 0) -> None:
    """Select the device (Neon::Backend with one device id)."""
    if device not in _initialised:
        check(lib.nx_init(device))
        _initialised.add(device)


def sync() -> None:
    check(lib.nx_sync())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib.nx_device_name(buf, 256))
    return buf.value.decode()


def spheres(dim, n: int, rmin: float, rmax: float, seed: int, periodic: int = 0) -> np.ndarray:
    """Solid mask (uint8, x fastest) of n seeded random spheres
    (nx_geometry_spheres)."""
    d = (C.c_int32 * 3)(*[int(v) for v in dim])
    out = np.empty(int(dim[0]) * int(dim[1]) * int(dim[2]), np.uint8)
    check(lib.nx_geometry_spheres(d, int(n), float(rmin), float(rmax), int(seed), int(periodic),
                                  out.ctypes.data))
    return out


def cavity_classes(dim, n_boxes: int = 0, emin: int = 4, emax: int = 16, seed: int = 0) -> np.ndarray:
    """Cavity classification (uint8 NX_CELL_*, x fastest) with seeded
    bounce-back boxes (nx_geometry_cavity)."""
    d = (C.c_int32 * 3)(*[int(v) for v in dim])
    out = np.empty(int(dim[0]) * int(dim[1]) * int(dim[2]), np.uint8)
    check(lib.nx_geometry_cavity(d, int(n_boxes), int(emin), int(emax), int(seed),
                                 out.ctypes.data))
    return out


def batch_cavities(n: int = 512, seed: int = 2025, edges=tuple(range(32, 97, 8)),
                   re_range=(50.0, 1000.0), boxes=(0, 4), box_edges=(4, 16), ulb: float = 0.05):
    """The batched-cavity workload: n lid-driven cubes, edge drawn from
    `edges`, Re uniform in re_range, boxes[0]..boxes[1] seeded bounce-back
    boxes with edges uniform in box_edges, lid ulb along x, nu = ulb * edge /
    Re.  One generator draws (edge, Re, box count, box seed) per member in
    that order.  Returns member dicts for CavityBatch (plus "re")."""
    rng = np.random.default_rng(seed)
    members = []
    for _ in range(int(n)):
        e = int(rng.choice(np.asarray(edges)))
        re = float(rng.uniform(*re_range))
        nb = int(rng.integers(boxes[0], boxes[1] + 1))
        cls = cavity_classes((e, e, e), nb, box_edges[0], box_edges[1], seed=int(rng.integers(0, 2 ** 63)))
        nu = ulb * e / re
        members.append({"dim": (e, e, e), "classes": cls, "omega": 1.0 / (3.0 * nu + 0.5),
                        "ulb": ulb, "ulid": (1.0, 0.0, 0.0), "re": re, "boxes": nb})
    return members


class DeviceBuffer:
    """Owning device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.nx_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        check(lib.nx_h2d(self.ptr, a.ctypes.data, a.nbytes, stream))
        check(lib.nx_stream_sync(stream))

    def download(self, dtype, count: int | None = None, stream=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dtype)
        check(lib.nx_d2h(out.ctypes.data, self.ptr, out.nbytes, stream))
        check(lib.nx_stream_sync(stream))
        return out

    def zero(self, stream=None) -> None:
        check(lib.nx_memset(self.ptr, 0, self.nbytes, stream))

    def free(self) -> None:
        if self.ptr:
            lib.nx_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib.nx_stream_create(C.byref(s)))
        self.handle = s.value

    def sync(self) -> None:
        check(lib.nx_stream_sync(self.handle))

    def __del__(self):
        try:
            lib.nx_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib.nx_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None) -> None:
        check(lib.nx_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def sync(self) -> None:
        check(lib.nx_event_sync(self.handle))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(lib.nx_event_elapsed_ms(self.handle, later.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            lib.nx_event_destroy(self.handle)
        except Exception:
            pass


def params(omega: float, force=(0.0, 0.0, 0.0)) -> abi.LbmParams:
    p = abi.LbmParams()
    p.omega = float(omega)
    for a in range(3):
        p.force[a] = float(force[a])
    return p


def _storage(s) -> int:
    return abi.STORAGE_NAMES[s] if isinstance(s, str) else int(s)


def _compute(c) -> int:
    return abi.COMPUTE_NAMES[c] if isinstance(c, str) else int(c)


def default_compute(storage: int) -> int:
    return abi.NX_COMPUTE_F64 if storage == abi.NX_F64 else abi.NX_COMPUTE_F32


class _Lattice:
    """State shared by the dense and block lattices (LbmIterationD3Q19)."""

    storage: int
    compute: int
    n_slots: int  # cells (dense) or voxel slots (block)

    def _alloc(self):
        nbytes = self._pop_bytes()
        self.pop = [DeviceBuffer(nbytes), DeviceBuffer(nbytes)]
        self.flags = DeviceBuffer(4 * self.n_slots)
        self.parity = 0
        self._out = DeviceBuffer(32)

    # which field holds the current populations (LbmIteration.h:58-69)
    @property
    def fin(self) -> DeviceBuffer:
        return self.pop[self.parity]

    @property
    def fout(self) -> DeviceBuffer:
        return self.pop[self.parity ^ 1]

    @property
    def np_storage(self):
        return _NP_STORAGE[self.storage]

    @property
    def moment_dtype(self):
        return np.float64 if self.storage == abi.NX_F64 else np.float32

    def populations(self) -> np.ndarray:
        """Raw stored populations of the current field (storage encoding)."""
        return self.fin.download(self.np_storage, 19 * self.n_slots)

    def set_populations(self, arr: np.ndarray, which: int | None = None) -> None:
        arr = np.ascontiguousarray(arr, dtype=self.np_storage).reshape(-1)
        assert arr.size == 19 * self.n_slots
        drop = getattr(self, "_drop_containers", None)
        if drop is not None:
            drop()                  # containers hold what they found on the walls
        if which is None:
            for b in self.pop:
                b.upload(arr)
        else:
            self.pop[which].upload(arr)

    def get_flags(self) -> np.ndarray:
        return self.flags.download(np.uint32, self.n_slots)

    def run(self, nsteps: int, omega: float, force=(0.0, 0.0, 0.0), stream=None) -> None:
        p = params(omega, force)
        for _ in range(int(nsteps)):
            self._step(p, stream)
            self.parity ^= 1

    def mass(self, stream=None) -> float:
        self._mass(stream)
        return float(self._out.download(np.float64, 1, stream)[0])

    def residual(self, u_prev: DeviceBuffer, stream=None) -> tuple[float, float]:
        self._residual(u_prev, stream)
        r = self._out.download(np.float64, 2, stream)
        return float(r[0]), float(r[1])


class DenseLattice(_Lattice):
    """D3Q19 on a dense grid (one dGrid partition), SoA layout."""

    def __init__(self, dim, storage="f32", compute=None, periodic: int = 0):
        init()
        self.dim = tuple(int(v) for v in dim)
        self.storage = _storage(storage)
        self.compute = default_compute(self.storage) if compute is None else _compute(compute)
        self.periodic = int(periodic)
        self.desc = abi.DenseDesc()
        for a in range(3):
            self.desc.dim[a] = self.dim[a]
        self.desc.storage, self.desc.compute, self.desc.periodic = (
            self.storage, self.compute, self.periodic)
        self.n_slots = self.dim[0] * self.dim[1] * self.dim[2]
        self._alloc()

    def _pop_bytes(self) -> int:
        return int(lib.nx_pop_bytes_dense(C.byref(self.desc)))

    def setup_cavity(self, ulb: float, ulid=(1.0, 0.0, 0.0), stream=None) -> None:
        """RunCavityTwoPop.cu:159-242: classification, populations, wall mask."""
        check(lib.nx_cavity_flags_dense(C.byref(self.desc), self.flags.ptr, stream))
        self._fill_and_mask(ulb, ulid, stream)

    def set_classes(self, cls: np.ndarray, ulb: float = 0.0, ulid=(1.0, 0.0, 0.0),
                    stream=None, zouhe: np.ndarray | None = None) -> None:
        """Arbitrary classification (x fastest), then fill + wall mask.
        zouhe: optional per-cell Zou-He code for bulk cells (face
        NX_ZH_XMIN..NX_ZH_ZMAX, plus 8 for a prescribed velocity), stepped by
        run_zouhe."""
        cls = np.ascontiguousarray(cls, dtype=np.uint32).reshape(-1)
        assert cls.size == self.n_slots
        fl = cls << abi.NX_CLASS_SHIFT
        if zouhe is not None:
            zh = np.ascontiguousarray(zouhe, dtype=np.uint32).reshape(-1)
            assert zh.size == self.n_slots and (zh & 7).max() <= 6 and zh.max() < 16
            fl = fl | (zh << abi.NX_ZH_SHIFT)
        self.flags.upload(fl)
        self._fill_and_mask(ulb, ulid, stream)

    def run_zouhe(self, nsteps: int, omega: float, values, stream=None) -> None:
        """Steps through nx_d3q19_step_dense_zouhe: values[face - 1] is the
        prescribed density (or inward normal velocity) of that face."""
        v = (C.c_double * 6)(*[float(x) for x in values])
        p = params(omega)
        for _ in range(int(nsteps)):
            check(lib.nx_d3q19_step_dense_zouhe(C.byref(self.desc), self.fin.ptr, self.fout.ptr,
                                                self.flags.ptr, C.byref(p), v, stream))
            self.parity ^= 1

    def _fill_and_mask(self, ulb, ulid, stream):
        self._drop_containers()
        lid = (C.c_double * 3)(*[float(v) for v in ulid])
        for b in self.pop:
            check(lib.nx_d3q19_fill_dense(C.byref(self.desc), b.ptr, self.flags.ptr, float(ulb),
                                          lid, stream))
        check(lib.nx_d3q19_wall_mask_dense(C.byref(self.desc), self.flags.ptr, self.flags.ptr,
                                           stream))
        check(lib.nx_stream_sync(stream))
        self.parity = 0

    def _step(self, p, stream):
        check(lib.nx_d3q19_step_dense(C.byref(self.desc), self.fin.ptr, self.fout.ptr,
                                      self.flags.ptr, C.byref(p), stream))

    def moments(self, stream=None):
        n = self.n_slots
        it = np.dtype(self.moment_dtype).itemsize
        rho, u = DeviceBuffer(n * it), DeviceBuffer(3 * n * it)
        check(lib.nx_d3q19_moments_dense(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                         rho.ptr, u.ptr, stream))
        X, Y, Z = self.dim
        r = rho.download(self.moment_dtype, n, stream).reshape(Z, Y, X)
        uu = u.download(self.moment_dtype, 3 * n, stream).reshape(3, Z, Y, X)
        return r, uu

    def tuning(self) -> dict:
        t = abi.Tuning()
        check(lib.nx_d3q19_get_tuning_dense(C.byref(self.desc), C.byref(t)))
        return t.as_dict()

    def set_tuning(self, wg=256, cpt=1, load_policy=0, xcd_remap=0, vx=1, boundary=0) -> None:
        t = abi.Tuning(int(wg), int(cpt), int(load_policy), int(xcd_remap), int(vx), int(boundary))
        check(lib.nx_d3q19_set_tuning_dense(C.byref(self.desc), C.byref(t)))

    # -- containers (LbmIterationD3Q19: one per ping-pong direction) ---------
    def _containers(self, omega, force):
        key = (float(omega), tuple(float(v) for v in force))
        if getattr(self, "_ckey", None) != key:
            self._drop_containers()
            p = params(omega, force)
            cs = []
            for k in range(2):
                h = C.c_void_p()
                check(lib.nx_d3q19_iteration_dense(C.byref(self.desc), self.pop[k].ptr,
                                                   self.pop[k ^ 1].ptr, self.flags.ptr,
                                                   C.byref(p), C.byref(h)))
                cs.append(h.value)
            self._cont, self._ckey = cs, key
        return self._cont

    def _drop_containers(self):
        for h in getattr(self, "_cont", []) or []:
            lib.nx_container_destroy(h)
        self._cont, self._ckey = [], None

    def container_info(self, omega, force=(0.0, 0.0, 0.0)):
        """(interior cells, other bulk cells, walls copied) of the container
        that runs the next step."""
        a, b, w = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib.nx_container_info(self._containers(omega, force)[self.parity], C.byref(a),
                                    C.byref(b), C.byref(w)))
        return int(a.value), int(b.value), bool(w.value)

    def run(self, nsteps: int, omega: float, force=(0.0, 0.0, 0.0), stream=None) -> None:
        """nsteps container steps in one call (hipGraph replays on a created
        stream, LbmIterationD3Q19::run)."""
        cs = self._containers(omega, force)
        n = int(nsteps)
        check(lib.nx_container_run_pair(cs[self.parity], cs[self.parity ^ 1], n, stream))
        self.parity ^= n & 1

    def run_plain(self, nsteps: int, omega: float, force=(0.0, 0.0, 0.0), stream=None) -> None:
        """Steps through nx_d3q19_step_dense (no container)."""
        _Lattice.run(self, nsteps, omega, force, stream)

    def autotune(self, omega: float, force=(0.0, 0.0, 0.0), stream=None):
        """Time every launch configuration on this lattice's buffers and keep
        the fastest (the output field is overwritten)."""
        t, ms = abi.Tuning(), C.c_float()
        h = self._containers(omega, force)[self.parity]
        check(lib.nx_container_autotune(h, stream, C.byref(t), C.byref(ms)))
        return t.as_dict(), float(ms.value)

    def __del__(self):
        try:
            self._drop_containers()
        except Exception:
            pass

    def _mass(self, stream):
        check(lib.nx_d3q19_mass_dense(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                      self._out.ptr, stream))

    def _residual(self, u_prev, stream):
        check(lib.nx_d3q19_residual_dense(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                          u_prev.ptr, self._out.ptr, stream))


class BlockGrid:
    """Host bGrid structure (nx_bgrid_create) plus its device tables."""

    def __init__(self, dim, active: np.ndarray, periodic: int = 0):
        init()
        self.dim = tuple(int(v) for v in dim)
        act = np.ascontiguousarray(active, dtype=np.uint8).reshape(-1)
        assert act.size == self.dim[0] * self.dim[1] * self.dim[2]
        d = (C.c_int32 * 3)(*self.dim)
        h = C.c_void_p()
        check(lib.nx_bgrid_create(d, act.ctypes.data, int(periodic), C.byref(h)))
        try:
            nb, na = C.c_uint32(), C.c_uint64()
            check(lib.nx_bgrid_info(h, C.byref(nb), C.byref(na)))
            self.n_blocks, self.n_active = int(nb.value), int(na.value)
            self.origins = np.empty(3 * self.n_blocks, np.int32)
            self.masks = np.empty(16 * self.n_blocks, np.uint32)
            self.conn = np.empty(27 * self.n_blocks, np.uint32)
            check(lib.nx_bgrid_tables(h, self.origins.ctypes.data, self.masks.ctypes.data,
                                      self.conn.ctypes.data))
        finally:
            lib.nx_bgrid_destroy(h)
        self.periodic = int(periodic)
        self.d_origins = DeviceBuffer.from_array(self.origins)
        self.d_masks = DeviceBuffer.from_array(self.masks)
        self.d_conn = DeviceBuffer.from_array(self.conn)


class BlockLattice(_Lattice):
    """D3Q19 on the block-sparse grid (one bGrid<8,8,8> partition), AoSoA layout."""

    def __init__(self, grid: BlockGrid, storage="f32", compute=None):
        self.grid = grid
        self.dim = grid.dim
        self.storage = _storage(storage)
        self.compute = default_compute(self.storage) if compute is None else _compute(compute)
        self.desc = abi.BlockDesc()
        for a in range(3):
            self.desc.dim[a] = self.dim[a]
        self.desc.storage, self.desc.compute, self.desc.periodic = (
            self.storage, self.compute, grid.periodic)
        self.desc.n_blocks = grid.n_blocks
        self.desc.origins, self.desc.masks, self.desc.conn = (
            grid.d_origins.ptr, grid.d_masks.ptr, grid.d_conn.ptr)
        self.n_slots = grid.n_blocks * 512
        self._tiles = None
        self._alloc()

    def _pop_bytes(self) -> int:
        return int(lib.nx_pop_bytes_block(C.byref(self.desc)))

    def set_tiled(self, on: bool = True) -> None:
        """Step with the block neighbourhood staged in LDS (one workgroup per
        2 x 2 x 1 tile of blocks, nx_d3q19_step_block_tiled) instead of one
        workgroup per block; both give the same populations bit for bit."""
        if on and self._tiles is None:
            h = C.c_void_p()
            check(lib.nx_block_tiles_create(C.byref(self.desc), C.byref(h)))
            self._tiles = h
        elif not on and self._tiles is not None:
            lib.nx_block_tiles_destroy(self._tiles)
            self._tiles = None

    @property
    def tiled(self) -> bool:
        return self._tiles is not None

    @property
    def n_tiles(self) -> int:
        if self._tiles is None:
            return 0
        n = C.c_uint32()
        check(lib.nx_block_tiles_info(self._tiles, C.byref(n)))
        return int(n.value)

    def __del__(self):
        try:
            if getattr(self, "_tiles", None) is not None:
                lib.nx_block_tiles_destroy(self._tiles)
                self._tiles = None
        except Exception:
            pass

    def setup_cavity(self, ulb: float, ulid=(1.0, 0.0, 0.0), stream=None) -> None:
        check(lib.nx_cavity_flags_block(C.byref(self.desc), self.flags.ptr, stream))
        self._fill_and_mask(ulb, ulid, stream)

    def set_classes(self, cls_slots: np.ndarray, ulb: float = 0.0, ulid=(1.0, 0.0, 0.0),
                    stream=None) -> None:
        """Classification per voxel slot (b*512 + v), then fill + wall mask."""
        cls = np.ascontiguousarray(cls_slots, dtype=np.uint32).reshape(-1)
        assert cls.size == self.n_slots
        self.flags.upload(cls << abi.NX_CLASS_SHIFT)
        self._fill_and_mask(ulb, ulid, stream)

    def _fill_and_mask(self, ulb, ulid, stream):
        lid = (C.c_double * 3)(*[float(v) for v in ulid])
        for b in self.pop:
            check(lib.nx_d3q19_fill_block(C.byref(self.desc), b.ptr, self.flags.ptr, float(ulb),
                                          lid, stream))
        check(lib.nx_d3q19_wall_mask_block(C.byref(self.desc), self.flags.ptr, self.flags.ptr,
                                           stream))
        check(lib.nx_stream_sync(stream))
        self.parity = 0

    def _step(self, p, stream):
        if self._tiles is not None:
            check(lib.nx_d3q19_step_block_tiled(C.byref(self.desc), self._tiles, self.fin.ptr,
                                                self.fout.ptr, self.flags.ptr, C.byref(p), stream))
            return
        check(lib.nx_d3q19_step_block(C.byref(self.desc), self.fin.ptr, self.fout.ptr,
                                      self.flags.ptr, C.byref(p), stream))

    def moments(self, stream=None):
        n = self.n_slots
        it = np.dtype(self.moment_dtype).itemsize
        rho, u = DeviceBuffer(n * it), DeviceBuffer(3 * n * it)
        check(lib.nx_d3q19_moments_block(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                         rho.ptr, u.ptr, stream))
        return rho.download(self.moment_dtype, n, stream), u.download(self.moment_dtype, 3 * n,
                                                                       stream)

    def _mass(self, stream):
        check(lib.nx_d3q19_mass_block(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                      self._out.ptr, stream))

    def _residual(self, u_prev, stream):
        check(lib.nx_d3q19_residual_block(C.byref(self.desc), self.fin.ptr, self.flags.ptr,
                                          u_prev.ptr, self._out.ptr, stream))


class CavityBatch:
    """Many independent dense cavities stepped by one persistent launch
    (nx_batch_*).  members: list of dicts with keys dim, classes (uint8
    NX_CELL_*), omega, ulb, ulid."""

    def __init__(self, members, storage="f32"):
        init()
        self.n = len(members)
        self.storage = _storage(storage)
        dims = np.array([m["dim"] for m in members], np.int32).reshape(-1)
        classes = np.concatenate([np.ascontiguousarray(m["classes"], np.uint8).reshape(-1)
                                  for m in members])
        omega = np.array([m["omega"] for m in members], np.float64)
        ulb = np.array([m.get("ulb", 0.05) for m in members], np.float64)
        ulid = np.array([m.get("ulid", (1.0, 0.0, 0.0)) for m in members], np.float64).reshape(-1)
        h = C.c_void_p()
        check(lib.nx_batch_create(self.n, dims.ctypes.data, classes.ctypes.data,
                                  omega.ctypes.data, ulb.ctypes.data, ulid.ctypes.data,
                                  self.storage, C.byref(h)))
        self.h = h.value
        self.dims = [tuple(int(v) for v in m["dim"]) for m in members]
        nm, nc, nt = C.c_uint32(), C.c_uint64(), C.c_uint32()
        check(lib.nx_batch_info(self.h, C.byref(nm), C.byref(nc), C.byref(nt)))
        self.n_cells, self.n_tiles = int(nc.value), int(nt.value)

    def tiles(self) -> int:
        """Tiles of the current launch (they depend on its workgroup size)."""
        nt = C.c_uint32()
        check(lib.nx_batch_info(self.h, None, None, C.byref(nt)))
        return int(nt.value)

    def init(self, rho=None, u=None, stream=None):
        """Equilibrium start from device buffers (arena cell order) or rest."""
        check(lib.nx_batch_init(self.h, rho.ptr if rho is not None else None,
                                u.ptr if u is not None else None, stream))

    def step(self, k: int, stream=None):
        check(lib.nx_batch_step(self.h, int(k), stream))

    def constrain(self, u, iterations: int, members=None, stream=None):
        """Consistent start after init(rho, u): `iterations` x (2 steps + move
        the equilibrium part back to the prescribed u), for the members whose
        entry of the boolean mask `members` is set (None: all)."""
        sel = None
        if members is not None:
            sel = np.ascontiguousarray(members, np.uint8)
            assert sel.size == self.n
        check(lib.nx_batch_constrain(self.h, u.ptr, int(iterations),
                                     sel.ctypes.data if sel is not None else None, stream))

    def run(self, max_steps: int, check_every: int, tol: float):
        steps = np.zeros(self.n, np.int32)
        res = np.zeros(self.n, np.float64)
        left = C.c_uint32()
        check(lib.nx_batch_run(self.h, int(max_steps), int(check_every), float(tol),
                               steps.ctypes.data, res.ctypes.data, C.byref(left)))
        return steps, res, int(left.value)

    def set_launch(self, wg: int, wgs_per_cu: int):
        check(lib.nx_batch_set_launch(self.h, int(wg), int(wgs_per_cu)))

    def set_variant(self, xcd_chunks: int = 1, edge: int = 1):
        check(lib.nx_batch_set_variant(self.h, int(xcd_chunks), int(edge)))

    def launch(self) -> dict:
        v = [C.c_int32() for _ in range(4)]
        check(lib.nx_batch_get_launch(self.h, *(C.byref(x) for x in v)))
        out = dict(zip(("wg", "wgs_per_cu", "xcd_chunks", "edge"), (int(x.value) for x in v)))
        grid, chunked = C.c_uint32(), C.c_int32()
        check(lib.nx_batch_launch_grid(self.h, C.byref(grid), C.byref(chunked)))
        out["grid"], out["chunked"] = int(grid.value), int(chunked.value)
        return out

    def autotune(self, stream=None):
        w, c, ms = C.c_int32(), C.c_int32(), C.c_float()
        check(lib.nx_batch_autotune(self.h, stream, C.byref(w), C.byref(c), C.byref(ms)))
        return self.launch(), float(ms.value)

    def member(self, k: int, which: int = 0):
        """(desc, populations pointer, flags pointer) of member k."""
        d, p, f = abi.DenseDesc(), C.c_void_p(), C.c_void_p()
        check(lib.nx_batch_member(self.h, int(k), int(which), C.byref(d), C.byref(p), C.byref(f)))
        return d, p.value, f.value

    def populations(self, k: int) -> np.ndarray:
        d, p, _ = self.member(k)
        n = 19 * d.dim[0] * d.dim[1] * d.dim[2]
        out = np.empty(n, _NP_STORAGE[self.storage])
        check(lib.nx_d2h(out.ctypes.data, p, out.nbytes, None))
        check(lib.nx_stream_sync(None))
        return out

    def mass(self, k: int) -> float:
        """FP64 mass of member k's latest populations (nx_d3q19_mass_dense on
        its dense view)."""
        d, p, f = self.member(k)
        out = DeviceBuffer(8)
        check(lib.nx_d3q19_mass_dense(C.byref(d), p, f, out.ptr, None))
        return float(out.download(np.float64, 1)[0])

    def moments(self, k: int):
        d, p, f = self.member(k)
        n = d.dim[0] * d.dim[1] * d.dim[2]
        mt = np.float64 if self.storage == abi.NX_F64 else np.float32
        it = np.dtype(mt).itemsize
        rho, u = DeviceBuffer(n * it), DeviceBuffer(3 * n * it)
        check(lib.nx_d3q19_moments_dense(C.byref(d), p, f, rho.ptr, u.ptr, None))
        X, Y, Z = d.dim[0], d.dim[1], d.dim[2]
        return (rho.download(mt, n).reshape(Z, Y, X), u.download(mt, 3 * n).reshape(3, Z, Y, X))

    def __del__(self):
        try:
            lib.nx_batch_destroy(self.h)
        except Exception:
            pass
