"""ctypes mirror of the C types in include/astaroth.h.

The enum tables are read from include/user_defines.h (the single source of their order, which
is ABI), so the Python names and values cannot drift from the C header.
Reference layout: include/astaroth.h:109-119 (AcMeshInfo, AcMesh).
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(REPO, "include")


def _read_table(macro: str) -> list[str]:
    with open(os.path.join(INCLUDE, "user_defines.h")) as fh:
        text = fh.read()
    m = re.search(r"#define\s+%s\(FUNC\)(.*?)(?:\n\s*\n|\n#|\s*\Z)" % re.escape(macro), text, re.S)
    if m is None:
        raise RuntimeError(f"table {macro} not found in user_defines.h")
    return re.findall(r"FUNC\((\w+)\)", m.group(1))


INT_PARAMS = _read_table("AC_FOR_USER_INT_PARAM_TYPES")
INT3_PARAMS = _read_table("AC_FOR_USER_INT3_PARAM_TYPES")
REAL_PARAMS = _read_table("AC_FOR_USER_REAL_PARAM_TYPES")
VTXBUF_NAMES = _read_table("AC_FOR_VTXBUF_HANDLES")
RTYPES = _read_table("AC_FOR_RTYPES")

NUM_INT_PARAMS = len(INT_PARAMS)
NUM_INT3_PARAMS = len(INT3_PARAMS)
NUM_REAL_PARAMS = len(REAL_PARAMS)
NUM_VTXBUF_HANDLES = len(VTXBUF_NAMES)
NGHOST = 3
STENCIL_ORDER = 6
NUM_STREAMS = 32
STREAM_DEFAULT = 0
STREAM_ALL = NUM_STREAMS

IP = {n: i for i, n in enumerate(INT_PARAMS)}
I3P = {n: i for i, n in enumerate(INT3_PARAMS)}
RP = {n: i for i, n in enumerate(REAL_PARAMS)}
VTX = {n: i for i, n in enumerate(VTXBUF_NAMES)}
RT = {n: i for i, n in enumerate(RTYPES)}

AC_SUCCESS = 0
AC_BOUNDCOND_PERIODIC, AC_BOUNDCOND_SYMMETRIC, AC_BOUNDCOND_ANTISYMMETRIC = 0, 1, 2
AC_FAILURE = 1


class Int3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class GridDims(ctypes.Structure):
    """GridDims of include/astaroth.h: m (with ghosts), n (interior)."""
    _fields_ = [("m", Int3), ("n", Int3)]


class DeviceConfiguration(ctypes.Structure):
    """DeviceConfiguration of include/astaroth.h (acNodeQueryDeviceConfiguration)."""
    _fields_ = [("num_devices", ctypes.c_int), ("devices", ctypes.POINTER(ctypes.c_void_p)),
                ("grid", GridDims), ("subgrid", GridDims)]


def int3(x, y, z) -> Int3:
    return Int3(int(x), int(y), int(z))


def real_ctype(double: bool):
    return ctypes.c_double if double else ctypes.c_float


def real_dtype(double: bool):
    return np.float64 if double else np.float32


_STRUCT_CACHE: dict[bool, tuple] = {}


def structs(double: bool):
    """(AcMeshInfo, AcMesh) ctypes classes for one precision."""
    if double in _STRUCT_CACHE:
        return _STRUCT_CACHE[double]
    R = real_ctype(double)

    class AcMeshInfo(ctypes.Structure):
        _fields_ = [
            ("int_params", ctypes.c_int * NUM_INT_PARAMS),
            ("int3_params", Int3 * NUM_INT3_PARAMS),
            ("real_params", R * NUM_REAL_PARAMS),
        ]

    class AcMesh(ctypes.Structure):
        _fields_ = [
            ("vertex_buffer", ctypes.POINTER(R) * NUM_VTXBUF_HANDLES),
            ("info", AcMeshInfo),
        ]

    _STRUCT_CACHE[double] = (AcMeshInfo, AcMesh)
    return AcMeshInfo, AcMesh


_SIMRES_CACHE = {}


def simulation_result(double: bool):
    """AcSimulationResult of include/astaroth_ext.h for one precision."""
    if double not in _SIMRES_CACHE:
        class AcSimulationResult(ctypes.Structure):
            _fields_ = [("last_step", ctypes.c_int), ("t", real_ctype(double)), ("stop", ctypes.c_int)]

        _SIMRES_CACHE[double] = AcSimulationResult
    return _SIMRES_CACHE[double]


STOP_REASONS = ("max_steps", "max_time", "small_dt", "nan", "stop_file")


class KernelConfig(ctypes.Structure):
    """AcKernelConfig of include/astaroth_ext.h."""

    _fields_ = [
        ("variant", ctypes.c_int),
        ("tile_x", ctypes.c_int),
        ("tile_y", ctypes.c_int),
        ("z_depth", ctypes.c_int),
    ]

    def as_tuple(self):
        return (self.variant, self.tile_x, self.tile_y, self.z_depth)

    def __repr__(self):
        return "KernelConfig(variant=%d, tile=(%d,%d), z_depth=%d)" % self.as_tuple()
