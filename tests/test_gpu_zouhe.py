"""Zou-He boundary cells on the GPU (nx_d3q19_step_dense_zouhe) against the
CPU oracle, bit for bit: a duct along each axis with Zou-He planes at both
ends (density / density and velocity / density), walls on the four other
faces, perturbed populations, binary64 and binary32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def duct(axis, dims, lo_code, hi_code):
    X, Y, Z = dims
    cls = np.full((Z, Y, X), 2, np.uint32)
    zh = np.zeros((Z, Y, X), np.uint32)
    idx = [slice(None)] * 3             # numpy order (z, y, x)
    npax = 2 - axis
    for ax in range(3):
        if ax == npax:
            continue
        s0 = [slice(None)] * 3
        s0[ax] = 0
        cls[tuple(s0)] = 0
        s0[ax] = -1
        cls[tuple(s0)] = 0
    lo = [slice(None)] * 3
    lo[npax] = 0
    hi = [slice(None)] * 3
    hi[npax] = -1
    zh[tuple(lo)] = lo_code
    zh[tuple(hi)] = hi_code
    zh[cls != 2] = 0
    return cls.reshape(-1), zh.reshape(-1)


CASES = [(0, 1, 2, [1.002, 0.998]), (0, 1 + 8, 2, [0.01, 1.0]), (1, 3, 4 + 8, [1.001, 0.02]),
         (2, 5 + 8, 6, [-0.015, 0.999]), (2, 5, 6, [0.997, 1.003])]


@pytest.mark.parametrize("st,cm,sto,cmo", [("f64", "f64", 0, 0), ("f32", "f32", 1, 1)])
@pytest.mark.parametrize("axis,lo,hi,vals", CASES)
def test_zouhe_duct_bit_exact(nx, orc, st, cm, sto, cmo, axis, lo, hi, vals):
    dims = (13, 10, 11)
    cls, zh = duct(axis, dims, lo, hi)
    g = nx.DenseLattice(dims, st, cm)
    o = orc.DenseOracle(dims, sto, cmo)
    g.set_classes(cls, zouhe=zh)
    o.set_classes(cls, zouhe=zh)
    assert np.array_equal(g.get_flags(), o.flags)
    assert ((o.flags >> orc.ZH_SHIFT) & 15).astype(bool).sum() == (zh > 0).sum()
    rng = np.random.default_rng(axis * 10 + lo)
    noise = 1.0 + 0.02 * rng.standard_normal(19 * cls.size)
    for k in (0, 1):
        p = (o.pop[k] * noise).astype(o.pop[k].dtype)
        o.pop[k][:] = p
        g.set_populations(p, which=k)
    v = [0.0] * 6
    v[(lo & 7) - 1] = vals[0]
    v[(hi & 7) - 1] = vals[1]
    g.run_zouhe(30, 1.4, v)
    o.run_zouhe(30, 1.4, v)
    assert np.array_equal(g.populations().view(np.uint8), o.fin.view(np.uint8))


def test_zouhe_cells_rejected_by_containers_and_deviation_storage(nx):
    dims = (12, 8, 8)
    cls, zh = duct(0, dims, 1, 2)
    g = nx.DenseLattice(dims, "f32", "f32")
    g.set_classes(cls, zouhe=zh)
    with pytest.raises(nx.abi.NeonError):
        g.run(1, 1.2)                        # the container refuses Zou-He cells
    import ctypes as C
    from neon_amd import abi
    p = nx.params(1.2, (1e-5, 0.0, 0.0))
    v = (C.c_double * 6)(1, 1, 0, 0, 0, 0)
    with pytest.raises(abi.NeonError):                # a body force is refused
        abi.check(abi.lib.nx_d3q19_step_dense_zouhe(C.byref(g.desc), g.fin.ptr, g.fout.ptr, g.flags.ptr,
                                                    C.byref(p), v, None))
    h = nx.DenseLattice(dims, "f32dev", "f32")
    h.set_classes(cls, zouhe=zh)
    with pytest.raises(nx.abi.NeonError):
        h.run_zouhe(1, 1.2, [1.0, 1.0, 0, 0, 0, 0])
