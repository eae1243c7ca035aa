"""The allocation-free residual (nx_d3q19_residual_dense_noalloc) against
nx_d3q19_residual_dense through the C ABI: the same sums, bit for bit, and
the same updated previous velocity, on two calls in a row, the second on a
smaller grid (fewer reduction partials than the first left in the static
scratch)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(nx, g, up_a, up_b):
    from neon_amd import abi
    out_a, out_b = nx.DeviceBuffer(16), nx.DeviceBuffer(16)
    abi.check(abi.lib.nx_d3q19_residual_dense(C.byref(g.desc), g.fin.ptr, g.flags.ptr, up_a.ptr,
                                              out_a.ptr, None))
    abi.check(abi.lib.nx_d3q19_residual_dense_noalloc(C.byref(g.desc), g.fin.ptr, g.flags.ptr,
                                                      up_b.ptr, out_b.ptr, None))
    abi.check(abi.lib.nx_sync())
    return out_a.download(np.float64, 2), out_b.download(np.float64, 2)


@pytest.mark.parametrize("dims,storage,omega", [
    ([(96, 80, 72), (12, 10, 9)], "f32", 1.7),
    ([(64, 64, 64), (20, 18, 16)], "f64", 1.6),
    ([(72, 40, 48), (16, 16, 16)], "f16dev", 1.8),
])
def test_noalloc_residual_equals_residual(nx, dims, storage, omega):
    for dim in dims:                       # the second grid is the smaller one
        g = nx.DenseLattice(dim, storage)
        g.setup_cavity(0.05)
        g.run(12, omega)
        n = dim[0] * dim[1] * dim[2]
        up_a, up_b = nx.DeviceBuffer(3 * n * 4), nx.DeviceBuffer(3 * n * 4)
        up_a.zero()
        up_b.zero()
        for rep in range(2):               # from u_prev = 0, then from the stored u
            a, b = _both(nx, g, up_a, up_b)
            assert a[1] > 0 and (rep == 0 or a[0] > 0), a
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (dim, rep, a, b)
            assert np.array_equal(up_a.download(np.uint32, 3 * n), up_b.download(np.uint32, 3 * n))
            g.run(4, omega)
