"""pagmo2_amd -- MI355X-native batch fitness evaluation for pagmo 2.x.

The product is libpagmo_hip.so (HIP kernels for gfx950 behind the C ABI of
include/pagmo_hip/pgh.h) and the C++ drop-in UDBFE include/pagmo_hip/hip_bfe.hpp.
This Python package only loads the library (pagmo2_amd._capi), builds it
(pagmo2_amd.build) and describes the uneven-archipelago workload
(pagmo2_amd.workloads).
"""
from . import _capi  # noqa: F401
from ._capi import *  # noqa: F401,F403
