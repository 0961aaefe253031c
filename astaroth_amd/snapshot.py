"""Readers of the snapshot files and time series written by acHostWriteMeshInfo,
acHostSaveMeshSnapshot / acDeviceSaveMeshSnapshot, acTimeseriesBegin / acDeviceTimeseriesAppend
and acDeviceRunSimulation.

The files are those of the reference's standalone driver (samples/standalone/simulation.cc:62-147,
251-362) and this module offers the entry points of the reference's analysis reader
(analysis/python/astar/data/read.py) under the same names, so analysis scripts written for it
can import this module instead:

    set_dtype(endian, AcRealSize)          read.py:33-42
    read_bin(fname, fdir, fnum, minfo)     read.py:44-66   -> (array[x, y, z], timestamp, ok)
    read_meshtxt(fdir, fname)              read.py:68-102  -> dict of every parameter
    parse_directory(meshdir)               read.py:104-112 -> sorted snapshot steps
    MeshInfo(fdir), Mesh(fnum, fdir)       read.py:251-299 (fields lnrho, ss, uu, aa, timestamp)
    parse_ts(fdir, fname), TimeSeries      read.py:612-662 (columns renamed as the reference does)

Plain numpy only; `fdir` is a path prefix as in the reference (a directory with its trailing
slash, or "" for the working directory).
"""
from __future__ import annotations

import os
import re

import numpy as np

FIELDS = ("VTXBUF_LNRHO", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ", "VTXBUF_AX", "VTXBUF_AY",
          "VTXBUF_AZ", "VTXBUF_ENTROPY")


def set_dtype(endian: int, AcRealSize: int, print_type: bool = False) -> np.dtype:
    """numpy dtype of a snapshot value: byte order from `endian` (1 little, 0 big), size in bytes."""
    dt = np.dtype(("<" if endian == 1 else ">") + "f" + str(int(AcRealSize)))
    if print_type:
        print("type_instruction", dt.str)
    return dt


def read_meshtxt(fdir: str, fname: str = "mesh_info.list", dbg_output: bool = False) -> dict:
    """The parameters of mesh_info.list: 'size_t'/'int' -> int, 'int3' -> [int]*3,
    'real' -> float, 'real3' -> [float]*3."""
    contents = {}
    with open(fdir + fname) as f:
        for raw in f:
            w = raw.split()
            if not w:
                continue
            kind, name, vals = w[0], w[1], w[2:]
            if kind in ("int", "size_t"):
                contents[name] = int(vals[0])
            elif kind == "int3":
                contents[name] = [int(v) for v in vals[:3]]
            elif kind == "real":
                contents[name] = float(vals[0])
            elif kind == "real3":
                contents[name] = [float(v) for v in vals[:3]]
            else:
                raise ValueError("mesh_info.list: unknown entry type %r" % kind)
            if dbg_output:
                print(name, contents[name])
    return contents


class MeshInfo:
    def __init__(self, fdir: str = "", dbg_output: bool = False):
        self.contents = read_meshtxt(fdir, "mesh_info.list", dbg_output)


def read_bin(fname: str, fdir: str, fnum, minfo: MeshInfo, numtype=None, getfilename: bool = False):
    """One field of snapshot `fnum`: (array indexed [x, y, z] over the whole m^3 box, time stamp,
    True), or (None, None, False) when the file does not exist."""
    path = "%s%s_%s.mesh" % (fdir, fname, fnum)
    if not os.path.exists(path):
        return None, None, False
    c = minfo.contents
    dt = set_dtype(c["endian"], c["AcRealSize"])
    if getfilename:
        print(path)
    raw = np.fromfile(path, dtype=dt)
    mx, my, mz = c["AC_mx"], c["AC_my"], c["AC_mz"]
    if raw.size != 1 + mx * my * mz:
        raise ValueError("%s holds %d values, expected 1 + %d" % (path, raw.size, mx * my * mz))
    # x fastest on disk, so Fortran order gives array[x, y, z]
    return np.reshape(raw[1:], (mx, my, mz), order="F"), raw[0], True


def parse_directory(meshdir: str) -> list:
    """Steps of the snapshots present in `meshdir` (by their VTXBUF_LNRHO files), sorted."""
    steps = []
    for name in os.listdir(meshdir):
        m = re.fullmatch(r"VTXBUF_LNRHO_(\d+)\.mesh", name)
        if m:
            steps.append(int(m.group(1)))
    return sorted(steps)


class Mesh:
    """The eight fields of one snapshot: lnrho, ss, uu = (ux, uy, uz), aa = (ax, ay, az), each
    indexed [x, y, z]; timestamp; ok (False when the snapshot is missing); grid coordinates."""

    def __init__(self, fnum, fdir: str = "", only_info: bool = False, pdiag: bool = False):
        fnum = str(fnum)
        self.framenum = fnum.zfill(10)
        self.minfo = MeshInfo(fdir)
        self.ok = False
        if only_info:
            return
        self.lnrho, self.timestamp, self.ok = read_bin("VTXBUF_LNRHO", fdir, fnum, self.minfo,
                                                       getfilename=pdiag)
        if not self.ok:
            return
        rd = lambda name: read_bin(name, fdir, fnum, self.minfo, getfilename=pdiag)[0]  # noqa: E731
        self.ss = rd("VTXBUF_ENTROPY")
        self.uu = (rd("VTXBUF_UUX"), rd("VTXBUF_UUY"), rd("VTXBUF_UUZ"))
        self.aa = (rd("VTXBUF_AX"), rd("VTXBUF_AY"), rd("VTXBUF_AZ"))
        c = self.minfo.contents
        self.xx = np.arange(c["AC_mx"]) * c["AC_dsx"]
        self.yy = np.arange(c["AC_my"]) * c["AC_dsy"]
        self.zz = np.arange(c["AC_mz"]) * c["AC_dsz"]
        self.xmid, self.ymid, self.zmid = c["AC_mx"] // 2, c["AC_my"] // 2, c["AC_mz"] // 2


def _ts_key(name: str) -> str:
    # the reference reader's column renaming, applied in the same order (read.py:623-635)
    for a, b in (("VTXBUF_", ""), ("UU", "uu"), ("_total", "tot"), ("ACCRETION", "acc"), ("A", "aa"),
                 ("LNRHO", "lnrho"), ("ENTROPY", "ss"), ("BFIELD", "bb"), ("X", "x"), ("Y", "y"),
                 ("Z", "z"), ("vaa", "vA")):
        name = name.replace(a, b)
    return name


def parse_ts(fdir: str = "", fname: str = "timeseries.ts", debug: bool = False) -> dict:
    """Columns of a time series by renamed title ('step', 't_step', 'dt', 'uutot_min', ...,
    'lnrho_rms', 'ss_max', ...), plus 'exist'. Rows are every complete data line (the reference
    reader drops the file's last line; a file written here ends each row with a line break, so
    nothing is dropped)."""
    path = fdir + fname
    var = {"exist": os.path.exists(path)}
    if not var["exist"]:
        return var
    with open(path) as f:
        lines = f.read().splitlines()
    title = lines[0].split()
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    rows = [r for r in rows if len(r) == len(title)]
    data = np.array(rows, dtype=np.float64).reshape(len(rows), len(title))
    for i, name in enumerate(title):
        var[_ts_key(name)] = data[:, i]
    var["step"] = var["step"].astype(np.int64)
    if debug:
        print(sorted(var))
    return var


class TimeSeries:
    def __init__(self, fdir: str = "", fname: str = "timeseries.ts", debug: bool = False):
        self.var = parse_ts(fdir, fname, debug=debug)
