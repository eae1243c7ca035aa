"""neon_amd -- MI355X-native D3Q19 lattice-Boltzmann step (dGrid / bGrid path).

The compute lives in libneon_amd.so (hand-written gfx950 HIP kernels behind
the C ABI of include/neon_amd.h).  This package only binds it.
"""
from . import abi  # noqa: F401  (raises ImportError when the library is not built)
from .lattice import (BlockGrid, BlockLattice, CavityBatch, DenseLattice, DeviceBuffer,  # noqa: F401
                      Event, Stream, batch_cavities, cavity_classes, device_name, init, params, spheres, sync)
