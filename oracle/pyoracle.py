"""ORACLE — test infrastructure, not product code.

ctypes bindings of oracle/liboracle_fp64.so (the CPU restatement of the reference host model
solver) for tests/ and the CPU-baseline leg of bench.py. Nothing in astaroth_amd/ imports this.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_SFX = {np.float32: "f32", np.float64: "f64", np.longdouble: "f80"}
_CT = {np.float32: ctypes.c_float, np.float64: ctypes.c_double, np.longdouble: ctypes.c_longdouble}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle_fp64.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib = ctypes.CDLL(path)
    return _lib


def _key(dtype):
    dt = np.dtype(dtype).type
    if dt not in _SFX:
        raise TypeError("unsupported oracle dtype %s" % dtype)
    return dt


def _ptrs(fields, dtype):
    arr = (ctypes.c_void_p * len(fields))()
    for q, f in enumerate(fields):
        assert f.dtype == dtype and f.flags["C_CONTIGUOUS"], (f.dtype, dtype)
        arr[q] = f.ctypes.data
    return arr


def _dims(fields):
    mz, my, mx = fields[0].shape
    return (ctypes.c_int * 3)(mx - 6, my - 6, mz - 6)


def _rp(real_params):
    rp = np.ascontiguousarray(real_params, dtype=np.float64)
    return rp, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def integrate_steps(fields, real_params, dt, nsteps=1):
    """acHostIntegrateStep applied nsteps times, in place (modelsolver.c:987-1041)."""
    t = _key(fields[0].dtype)
    fn = getattr(lib(), "oracle_integrate_steps_" + _SFX[t])
    rp, prp = _rp(real_params)
    fn(_dims(fields), _ptrs(fields, fields[0].dtype), prp, _CT[t](dt), int(nsteps))
    return fields


def integrate_steps_gpu_form(fields, real_params, dt, nsteps=1):
    """nsteps RK3 steps in the two-buffer form of the reference's GPU path (integration.cuh:14-37),
    driven as 3 x {periodic fill, substep, swap}; the new state is written back into `fields`."""
    t = _key(fields[0].dtype)
    fn = getattr(lib(), "oracle_integrate_steps_gpuform_" + _SFX[t])
    scratch = [np.zeros_like(f) for f in fields]
    rp, prp = _rp(real_params)
    fn(_dims(fields), _ptrs(fields, fields[0].dtype), _ptrs(scratch, fields[0].dtype), prp, _CT[t](dt), int(nsteps))
    return fields


def substep_gpu_form(in_fields, out_fields, real_params, step, dt, start, end):
    """acDeviceIntegrateSubstep semantics on the CPU: writes out_fields inside [start, end)."""
    t = _key(in_fields[0].dtype)
    fn = getattr(lib(), "oracle_substep_gpuform_" + _SFX[t])
    rp, prp = _rp(real_params)
    fn(_dims(in_fields), _ptrs(in_fields, in_fields[0].dtype), _ptrs(out_fields, in_fields[0].dtype), prp,
       int(step), _CT[t](dt), (ctypes.c_int * 3)(*start), (ctypes.c_int * 3)(*end))
    return out_fields


def rhs(fields, real_params):
    """Rate of change of the 8 fields at every interior cell (no ghost fill)."""
    t = _key(fields[0].dtype)
    rates = [np.zeros_like(f) for f in fields]
    fn = getattr(lib(), "oracle_rhs_" + _SFX[t])
    rp, prp = _rp(real_params)
    fn(_dims(fields), _ptrs(fields, fields[0].dtype), prp, _ptrs(rates, fields[0].dtype))
    return rates


def periodic(fields):
    """acHostMeshApplyPeriodicBounds, in place (memory.c:41-104)."""
    t = _key(fields[0].dtype)
    fn = getattr(lib(), "oracle_periodic_" + _SFX[t])
    fn(_dims(fields), _ptrs(fields, fields[0].dtype), len(fields))
    return fields


RTYPE = {"RTYPE_MAX": 0, "RTYPE_MIN": 1, "RTYPE_RMS": 2, "RTYPE_RMS_EXP": 3, "RTYPE_SUM": 7}


def reduce(fields_used, rtype):
    """acHostReduceScal (1 field) / acHostReduceVec (3 fields) (modelreduce.c:76-209)."""
    t = _key(fields_used[0].dtype)
    fn = getattr(lib(), "oracle_reduce_" + _SFX[t])
    out = _CT[t]()
    r = RTYPE[rtype] if isinstance(rtype, str) else int(rtype)
    res = fn(_dims(fields_used), r, len(fields_used), _ptrs(fields_used, fields_used[0].dtype), ctypes.byref(out))
    if res != 0:
        raise ValueError("unsupported rtype %s" % rtype)
    return t(out.value)


def derivs(field, inv_ds, i, j, k):
    """value, d/dx, d/dy, d/dz, d2/dx2, d2/dy2, d2/dz2, dxy, dxz, dyz at cell (i, j, k)."""
    t = _key(field.dtype)
    fn = getattr(lib(), "oracle_derivs_" + _SFX[t])
    out = np.zeros(10, dtype=field.dtype)
    inv = np.ascontiguousarray(inv_ds, dtype=np.float64)
    mz, my, mx = field.shape
    fn((ctypes.c_int * 3)(mx - 6, my - 6, mz - 6), ctypes.c_void_p(field.ctypes.data),
       inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(i), int(j), int(k),
       ctypes.c_void_p(out.ctypes.data))
    return out


# ---- acVerifyMesh-style comparison (verification.c:30-168), vectorised ------------------------
def ulp_error(model, candidate, p):
    """abs / ulp / rel error arrays as acGetError defines them (p = 24 or 53)."""
    m = np.asarray(model, dtype=np.longdouble)
    c = np.asarray(candidate, dtype=np.longdouble)
    absd = np.abs(m - c)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.floor(np.log2(np.abs(m)))
        ulp = np.power(np.longdouble(2), e - (p - 1))
        ulps = np.where(absd == 0, 0, absd / ulp)
        eps = np.longdouble(0.5) * np.power(np.longdouble(2), -(p - 1))
        rel = np.where(absd == 0, 0, np.abs(1 - c / m) / eps)
    return absd, ulps, rel


def verify_field(model, candidate, p, real_eps):
    """Errors at the point of maximum absolute error and the acEvalError verdict.

    Long-double inputs are compared in long double throughout. Otherwise the point is found on
    the FP64 differences (exact for values within a factor of two of each other, Sterbenz), the
    cells within 1e-6 of the largest one are compared again in long double, and the errors are
    evaluated there: the same point and values as the long-double path, at a fraction of its cost
    on large boxes."""
    model, candidate = np.asarray(model), np.asarray(candidate)
    if np.longdouble in (model.dtype.type, candidate.dtype.type):
        absd, ulps, rel = ulp_error(model, candidate, p)
        k = int(np.argmax(absd))
        abs_e, ulp_e, rel_e = float(absd.flat[k]), float(ulps.flat[k]), float(rel.flat[k])
    else:
        d = np.abs(model.astype(np.float64, copy=False) - candidate.astype(np.float64, copy=False))
        dmax = float(d.max())
        near = np.flatnonzero(d >= dmax * (1 - 1e-6)) if dmax > 0 else np.array([int(np.argmax(d))])
        del d
        absd, ulps, rel = ulp_error(model.flat[near], candidate.flat[near], p)
        j = int(np.argmax(absd))
        k = int(near[j])
        abs_e, ulp_e, rel_e = float(absd[j]), float(ulps[j]), float(rel[j])
    minmag = float(np.min(np.abs(model)))
    maxmag = float(np.max(np.abs(model)))
    ok = (ulp_e < 5) or (abs_e < minmag * real_eps)
    return dict(abs=abs_e, ulp=ulp_e, rel=rel_e, min=minmag, max=maxmag, ok=bool(ok),
                model=float(model.flat[k]), candidate=float(candidate.flat[k]))
