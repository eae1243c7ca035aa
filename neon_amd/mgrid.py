"""Grid refinement (Neon's mGrid, depth 2 to 5) driven through the C ABI
(nx_mgrid_*): the level activation of the reference's lid-driven cavity
(apps/lbmMultiRes/lidDrivenCavity.h:88-271), its --scale configurations, and
the solver handle."""
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


def level_dims(dim, depth: int):
    """Voxels per axis of every level: ceil(dim / 2^l) (a level-l voxel v is
    in the domain when its level-0 origin v 2^l is)."""
    out = [tuple(int(v) for v in dim)]
    for _ in range(1, depth):
        out.append(tuple((v + 1) // 2 for v in out[-1]))
    return out


def cavity_lambdas(dim, level_sdf=(0.0, -1.0 / 3.0, -1.0)):
    """Activation of the levels of the refined cavity: level l holds the
    voxels with level_sdf[l+1] < sdfCube <= level_sdf[l] (lidDrivenCavity.h:
    229-236; sdfCube of the level-0 position against dim - 1), one level per
    interval of level_sdf.  Returns a tuple with, for every level l, the
    lambda at every level-l voxel evaluated at its level-0 origin v 2^l,
    uint8, x fastest."""
    X, Y, Z = (int(v) for v in dim)
    box = np.array([X - 1, Y - 1, Z - 1])
    lv = [np.float32(v) for v in level_sdf]

    def lam(shape, scale, lo, hi):
        z, y, x = np.meshgrid(*(np.arange(n) for n in shape[::-1]), indexing="ij")
        ids = np.stack([x, y, z], -1) * scale
        s = sdf_cube(ids, box)
        return ((s <= lo) & (s > hi)).astype(np.uint8).reshape(-1)

    depth = len(lv) - 1
    return tuple(lam(d, 1 << l, lv[l], lv[l + 1]) for l, d in enumerate(level_dims(dim, depth)))


# The lid-driven cavity configurations of the multi-resolution app
# (apps/lbmMultiRes/lidDrivenCavity.h:95-214): --scale -> (grid edge,
# levelSDF); the depth is len(levelSDF) - 1.
LID_SCALES = {
    -1: (192, (0.0, -36.0 / 96.0, -72.0 / 96.0, -1.0)),
    0: (48, (0.0, -8.0 / 24.0, -16.0 / 24.0, -1.0)),
    1: (160, (0.0, -16.0 / 80.0, -32.0 / 80.0, -1.0)),
    2: (240, (0.0, -24.0 / 120.0, -80.0 / 120.0, -1.0)),
    3: (320, (0.0, -32.0 / 160.0, -64.0 / 160.0, -1.0)),
    4: (480, (0.0, -48.0 / 240.0, -96.0 / 240.0, -1.0)),
    5: (512, (0.0, -64.0 / 256.0, -112.0 / 256.0, -1.0)),
    6: (160, (0.0, -32.0 / 160.0, -64.0 / 160.0, -128.0 / 160.0, -1.0)),
    7: (240, (0.0, -32.0 / 120.0, -80.0 / 120.0, -112.0 / 120.0, -1.0)),
    8: (320, (0.0, -32.0 / 160.0, -64.0 / 160.0, -112.0 / 160.0, -1.0)),
    9: (480, (0.0, -48.0 / 240.0, -96.0 / 240.0, -160.0 / 240.0, -1.0)),
    10: (512, (0.0, -120.0 / 512.0, -200.0 / 512.0, -400.0 / 512.0, -1.0)),
    11: (512, (0.0, -120.0 / 512.0, -200.0 / 512.0, -400.0 / 512.0, -1.0)),
}


def sphere_problem(scale: int = 2):
    """The multi-resolution app's flow over a sphere (flowOverShape.h:115-190)
    at --scale s: a 136s x 96s x 136s channel, three levels (level 0 inside
    [40s, 96s) x [40s, 64s) x [40s, 96s), level 1 inside [24s, 112s) x
    [24s, 72s) x [24s, 112s), level 2 everywhere), a sphere of radius 8s at
    (52s, 52s, 68s).  Returns (dim, lambdas per level, classes per level,
    sphere (cx, cy, cz, r)): the class of every level-l voxel from its level-0
    origin (flowOverShape.h:56-83): inlet (moving wall with the inlet
    velocity) at x = 0, bounce-back inside the sphere and on the y and z
    faces (y = 0 or the level's last voxel, likewise z), bulk elsewhere;
    the x = X - 1 face is left open."""
    from .abi import NX_CELL_BULK, NX_CELL_BOUNCEBACK, NX_CELL_MOVINGWALL
    s_ = int(scale)
    dim = (136 * s_, 96 * s_, 136 * s_)
    sph = (52 * s_, 52 * s_, 68 * s_, 8 * s_)
    boxes = [((40, 96), (40, 64), (40, 96)), ((24, 112), (24, 72), (24, 112))]
    dims = level_dims(dim, 3)
    lams, cls = [], []
    for l, d in enumerate(dims):
        z, y, x = np.meshgrid(*(np.arange(n) << l for n in d[::-1]), indexing="ij")
        if l < 2:
            (x0, x1), (y0, y1), (z0, z1) = boxes[l]
            lam = (x >= x0 * s_) & (x < x1 * s_) & (y >= y0 * s_) & (y < y1 * s_) & (z >= z0 * s_) & (z < z1 * s_)
        else:
            lam = np.ones(x.shape, bool)
        lams.append(lam.astype(np.uint8).reshape(-1))
        c = np.full(x.shape, NX_CELL_BULK, np.uint8)
        c[x == 0] = NX_CELL_MOVINGWALL
        inside = (sph[0] - x) ** 2 + (sph[1] - y) ** 2 + (sph[2] - z) ** 2 < sph[3] ** 2
        c[inside] = NX_CELL_BOUNCEBACK
        wall = (y == 0) | (y == dim[1] - (1 << l)) | (z == 0) | (z == dim[2] - (1 << l))
        c[wall] = NX_CELL_BOUNCEBACK
        cls.append(c.reshape(-1))
    return dim, lams, cls, sph


def sphere_parameters(scale: int, re: float, storage: str = "f64", uin: float = 0.04):
    """(uin, omega0) of the flow over a sphere in the field type T
    (flowOverShape.h:142-147): clength = radius / 2^(depth-1) with depth 3,
    nu = uin clength / Re, omega0 = 1 / (3 nu + 1/2)."""
    T = np.float64 if storage == "f64" else np.float32
    u = T(uin)
    clength = T((8 * int(scale)) // 4)
    nu = T(u * clength / T(re))
    return float(u), float(T(1.0 / (3.0 * float(nu) + 0.5)))


def cavity_parameters(dim, re: float, storage: str = "f64", ulb: float = 0.04, depth: int = 2):
    """(ulb, omega0) of the refined cavity in the field type T
    (lidDrivenCavity.h:246-251): ulb = 0.04 in T, the characteristic length is
    the top level's dimension, grid.getDimension(depth - 1).x = its groups of
    2 times 2 = 2 ceil(X / 2^depth) (mGrid.cpp:684-687), nu = ulb * length /
    Re in T, omega0 = 1 / (3 nu + 1/2) evaluated in binary64 and stored in T."""
    T = np.float64 if storage == "f64" else np.float32
    u = T(ulb)
    length = T(2 * ((int(dim[0]) + (1 << depth) - 1) >> depth))
    nu = T(u * length / T(re))
    omega = T(1.0 / (3.0 * float(nu) + 0.5))
    return float(u), float(omega)


class MGrid:
    """Refined grid and its LBM fields on the device (nx_mgrid_*).
    MGrid(dim, [lam_0, .., lam_{D-1}], storage) or, for two levels,
    MGrid(dim, lam0, lam1, storage): lam_l is the activation of level l at
    every level-l voxel (evaluated at its level-0 origin v 2^l), uint8, x
    fastest.  lattice "bgk19" (BGK D3Q19) or "kbc27" (KBC D3Q27, the
    multi-resolution app's default)."""

    LATTICES = {"bgk19": 0, "kbc27": 1}

    def __init__(self, dim, lams, *rest, lattice="bgk19"):
        from .lattice import init, _storage
        if lattice not in self.LATTICES:
            raise ValueError(f"lattice {lattice!r}: one of {sorted(self.LATTICES)}")
        self.lattice = lattice
        self.Q = 27 if lattice == "kbc27" else 19
        if isinstance(lams, (list, tuple)):
            storage = rest[0] if rest else "f64"
        else:  # (dim, lam0, lam1[, storage])
            lams = [lams, rest[0]]
            storage = rest[1] if len(rest) > 1 else "f64"
        self.dim = tuple(int(v) for v in dim)
        self.depth = len(lams)
        self.dims = level_dims(self.dim, self.depth)
        self.dim1 = self.dims[1]
        self.storage = _storage(storage)
        lams = [np.ascontiguousarray(a, np.uint8).reshape(-1) for a in lams]
        for a, d in zip(lams, self.dims):
            if a.size != int(np.prod(d)):
                raise ValueError(f"activation array of {a.size} voxels for a level of {d}")
        h = C.c_void_p()
        ptrs = (C.c_void_p * self.depth)(*[a.ctypes.data for a in lams])
        check(lib.nx_mgrid_create_levels((C.c_int32 * 3)(*self.dim), self.depth, ptrs, self.storage,
                                         self.LATTICES[lattice], C.byref(h)))
        self.h = h.value
        self._init = init

    def __del__(self):
        try:
            lib.nx_mgrid_destroy(self.h)
        except Exception:
            pass

    def state(self, level=None):
        """State bytes of one level, or of every level when level is None:
        bit 0 active, bit 1 refined, bits 2..3 the cell class."""
        if level is None:
            return tuple(self.state(l) for l in range(self.depth))
        st = np.zeros(int(np.prod(self.dims[level])), np.uint8)
        check(lib.nx_mgrid_state_level(self.h, int(level), st.ctypes.data))
        return st

    def info(self) -> dict:
        D = self.depth
        nb = (C.c_uint32 * D)()
        na, nl = (C.c_uint64 * D)(), (C.c_uint64 * D)()
        ns, fused = C.c_uint64(), C.c_int32()
        check(lib.nx_mgrid_info(self.h, nb, na, nl, C.byref(ns), C.byref(fused)))
        return {"blocks": list(nb), "active": list(na), "leaf": list(nl), "store_links": int(ns.value),
                "fine_fused": bool(fused.value & 1),
                "fused_levels": [bool((fused.value >> l) & 1) for l in range(D)]}

    def setup_cavity(self, ulb: float) -> None:
        self._init()
        check(lib.nx_mgrid_cavity(self.h, float(ulb)))

    def setup_cells(self, cls, u_wall) -> None:
        """Set-up with the class of every leaf per level (cls[l]: uint8 per
        level-l voxel, NX_CELL_BULK / BOUNCEBACK / MOVINGWALL) and the
        moving-wall velocity (lid or inlet)."""
        self._init()
        cls = [np.ascontiguousarray(c, np.uint8).reshape(-1) for c in cls]
        for c, d in zip(cls, self.dims):
            if c.size != int(np.prod(d)):
                raise ValueError(f"class array of {c.size} voxels for a level of {d}")
        ptrs = (C.c_void_p * self.depth)(*[c.ctypes.data for c in cls])
        check(lib.nx_mgrid_setup_cells(self.h, ptrs, (C.c_double * 3)(*[float(v) for v in u_wall])))

    def run(self, n: int, omega0: float, stream=None) -> None:
        check(lib.nx_mgrid_step(self.h, float(omega0), int(n), stream))

    def fields(self, level: int, which: int = 0) -> np.ndarray:
        """fin (which 0) or fout (which 1) of one level, dense [Q * n]."""
        n = int(np.prod(self.dims[level]))
        dt = np.float64 if self.storage == abi.NX_F64 else np.float32
        out = np.zeros(self.Q * n, dt)
        check(lib.nx_mgrid_fields(self.h, int(level), int(which), out.ctypes.data))
        return out
