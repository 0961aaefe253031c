"""acDeviceReduceNoAlloc (include/astaroth_ext.h) against the reductions it varies,
acDeviceReduce{Scal,Vec,VecScal}Double, through the C ABI: the same FP64 result bit for bit for
every reduction type, on two Devices in a row, the second one smaller (the library's per-GPU
scratch array keeps the first call's partials, which the second must not see). Bad arguments are
rejected on the host.
"""
import ctypes

import numpy as np
import pytest

from astaroth_amd import abi
from astaroth_amd.api import Library
from astaroth_amd.hiprt import Hip
from astaroth_amd.initial import device_init_mhd

pytestmark = pytest.mark.gpu

GROUPS = [
    (("VTXBUF_LNRHO",), ("RTYPE_MAX", "RTYPE_MIN", "RTYPE_RMS", "RTYPE_RMS_EXP", "RTYPE_SUM")),
    (("VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ"), ("RTYPE_MAX", "RTYPE_MIN", "RTYPE_RMS", "RTYPE_RMS_EXP",
                                                  "RTYPE_SUM")),
    (("VTXBUF_AX", "VTXBUF_AY", "VTXBUF_AZ", "VTXBUF_LNRHO"), ("RTYPE_ALFVEN_MAX", "RTYPE_ALFVEN_MIN",
                                                              "RTYPE_ALFVEN_RMS")),
]
SIZES = [(96, 80, 72), (20, 12, 9)]  # the second call's box is smaller


def existing(dev, rtype, names):
    if len(names) == 1:
        return dev.reduce_scal(rtype, names[0], double=True)
    if len(names) == 3:
        return dev.reduce_vec(rtype, *names, double=True)
    return dev.reduce_vec_scal(rtype, *names, double=True)


def no_alloc(lib, hip, dev, d_out, rtype, names):
    hip.memset(d_out, 0xFF, 8)  # a NaN pattern: the call must overwrite it
    handles = (ctypes.c_int * len(names))(*[abi.VTX[n] for n in names])
    r = lib.dll.acDeviceReduceNoAlloc(dev.handle, abi.STREAM_DEFAULT, abi.RT[rtype], len(names), handles,
                                      ctypes.c_void_p(d_out))
    assert r == abi.AC_SUCCESS
    dev.synchronize()
    return float(hip.to_host(d_out, 1, np.float64)[0])


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_no_alloc_reduction_equals_device_reduction(double):
    lib, hip = Library(double), Hip()
    d_out = hip.malloc(8)
    try:
        for dims in SIZES:
            info = lib.config(n=dims[0], ny=dims[1], nz=dims[2])
            with lib.device_create(0, info, autotune=False) as dev:
                device_init_mhd(dev, config_id=7)
                for names, rtypes in GROUPS:
                    for rt in rtypes:
                        want = existing(dev, rt, names)
                        got = no_alloc(lib, hip, dev, d_out, rt, names)
                        print("%s %s %s %-16s %.17g" % ("fp64" if double else "fp32", dims, len(names), rt, got),
                              flush=True)
                        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (dims, names, rt, got, want)
    finally:
        hip.free(d_out)


def test_no_alloc_reduction_rejects_bad_arguments():
    lib, hip = Library(False), Hip()
    d_out = hip.malloc(8)
    try:
        with lib.device_create(0, lib.config(n=16), autotune=False) as dev:
            f = lib.dll.acDeviceReduceNoAlloc
            one = (ctypes.c_int * 1)(abi.VTX["VTXBUF_LNRHO"])
            two = (ctypes.c_int * 2)(0, 1)
            bad = (ctypes.c_int * 1)(8)
            out = ctypes.c_void_p(d_out)
            assert f(dev.handle, 0, abi.RT["RTYPE_MAX"], 2, two, out) == abi.AC_FAILURE
            assert f(dev.handle, 0, abi.RT["RTYPE_ALFVEN_MAX"], 1, one, out) == abi.AC_FAILURE
            assert f(dev.handle, 0, abi.RT["RTYPE_MAX"], 1, bad, out) == abi.AC_FAILURE
            assert f(dev.handle, 0, abi.RT["RTYPE_MAX"], 1, one, None) == abi.AC_FAILURE
            assert f(dev.handle, abi.NUM_STREAMS, abi.RT["RTYPE_MAX"], 1, one, out) == abi.AC_FAILURE
            assert f(dev.handle, 0, abi.RT["RTYPE_MAX"], 1, one, out) == abi.AC_SUCCESS
            dev.synchronize()
    finally:
        hip.free(d_out)
