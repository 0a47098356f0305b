"""occ_amd: MI355X-native Kohn-Sham exchange-correlation build for OCC.

The compute path is liboccxc.so (HIP kernels for gfx950 behind the C ABI in
include/occxc.h); this package is the host-side mirror of OCC's interface for
that path: basis loading, molecular grids, and the DFT XC entry points.
"""
from .basis import AOBasis, Shell  # noqa: F401
from .molecules import water, water_cluster, alkane  # noqa: F401

__version__ = "0.1.0"
