"""GIO state files (src/cdg/io/gio.ipp) through gf_gio_write / gf_gio_read:
the header GIO::read_header reads (ivers, dim, nelems, porder, gtype, cycle,
time, multivar, element keys) followed by the field doubles."""
import ctypes

import numpy as np

from ._capi import GioInfo, check, lib

_dp = ctypes.POINTER(ctypes.c_double)


def write_state(path, fields, porder, gtype, nelems, time=0.0, cycle=0, ivers=0, keys=None):
    """Write one field (or several, multivar) of a grid with nelems elements;
    porder: one row of dim orders (ivers 0) or one row per element."""
    fields = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    po = np.ascontiguousarray(np.atleast_2d(porder), dtype=np.int32)
    info = GioInfo(ivers=ivers, dim=po.shape[1], nelems=nelems, gtype=gtype, cycle=cycle,
                   time=time, multivar=int(len(fields) > 1), nporder=po.size)
    kp = None
    if keys is not None:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        kp = keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    fp = (_dp * len(fields))(*[f.ctypes.data_as(_dp) for f in fields])
    check(lib().gf_gio_write(str(path).encode(), ctypes.byref(info),
                             po.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), kp, fp, len(fields),
                             fields[0].size))


def read_state(path, nfields=None):
    """Returns (info dict, porder rows, keys, list of fields)."""
    info = GioInfo()
    check(lib().gf_gio_read(str(path).encode(), ctypes.byref(info), None, 0, None, 0, None, 0, 0))
    po = np.zeros(info.nporder, dtype=np.int32)
    keys = np.zeros(info.nelems, dtype=np.uint64)
    check(lib().gf_gio_read(str(path).encode(), ctypes.byref(info),
                            po.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), po.size,
                            keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), keys.size,
                            None, 0, 0))
    rows = po.reshape(-1, info.dim)
    npts = int(np.sum(np.prod(rows + 1, axis=1))) if info.ivers else \
        int(info.nelems) * int(np.prod(rows[0] + 1))
    nf = nfields if nfields is not None else 1
    fields = [np.zeros(npts) for _ in range(nf)]
    fp = (_dp * nf)(*[f.ctypes.data_as(_dp) for f in fields])
    check(lib().gf_gio_read(str(path).encode(), ctypes.byref(info), None, 0, None, 0, fp, nf, npts))
    meta = dict(ivers=info.ivers, dim=info.dim, nelems=info.nelems, gtype=info.gtype,
                cycle=info.cycle, time=info.time, multivar=info.multivar)
    return meta, rows, keys, fields
