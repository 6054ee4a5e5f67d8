"""mesh_amd: MI355X-native lattice gauge hot path (SU(3) Wilson HMC, Landau
gauge fixing) behind a C ABI (include/mesh_hip.h, libmesh_hip.so).

The Python modules are a thin host-side mirror of krox/mesh's C++ call
surface; every computation runs in hand-written HIP kernels for gfx950.
"""
from ._lib import F32, F64, MeshError  # noqa: F401
from .lattice import Context, DeviceArray, Field, GaugeField, Grid, Lattice, default_context  # noqa: F401
from . import gauge  # noqa: F401
