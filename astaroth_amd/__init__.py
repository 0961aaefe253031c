"""MI355X-native single-device Astaroth stencil path.

The product is the C ABI of include/astaroth.h, implemented by the HIP libraries
astaroth_amd/lib/libastaroth_{fp32,fp64}.so (sources in astaroth_amd/csrc). This package holds
the build recipe (build.py), the ctypes mirror of the C interface (api.py, abi.py) and the
synthetic initial conditions of the benchmark configurations (initial.py).
"""
from .abi import IP, RP, VTX, RT, NGHOST  # noqa: F401
from .api import Library, Device, HostMesh, AstarothError, lib_path  # noqa: F401
