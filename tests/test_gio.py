"""GIO state files (src/cdg/io/gio.ipp) on the CPU.

The file is parsed here with `struct`, independently of the library, in the
field order of GIO::read_header (gio.ipp:588-640), and its size is checked
against GIO::sz_header (gio.ipp:678-690): (rows*dim + 4) ints + 2 size_t +
1 double + nelems keys, then the doubles.  Round trips cover ivers 0 and 1,
one field and multivar files, and the element keys.

Not kept: GIO::write_posix (gio.ipp:346-372) writes the header, closes the
file and reopens it with "wb" for the data, which truncates the header away;
the collective writer and the reader both expect header + data, which is
what gf_gio_write produces.  analysis/python/g2vtk (read_header, lines
158-196) reads an older header without multivar and keys.
"""
import struct

import numpy as np

import geoflow_amd as ga
from geoflow_amd.gio import read_state, write_state


def _parse(path):
    b = open(path, "rb").read()
    off = 0

    def take(fmt, n=1):
        nonlocal off
        v = struct.unpack_from("@" + fmt * n, b, off)
        off += struct.calcsize("@" + fmt * n)
        return v if n > 1 else v[0]

    h = {}
    h["ivers"], h["dim"] = take("i"), take("i")
    h["nelems"] = take("Q")
    rows = 1 if h["ivers"] == 0 else h["nelems"]
    po = take("i", rows * h["dim"])
    h["porder"] = np.reshape(po if rows * h["dim"] > 1 else [po], (rows, h["dim"]))
    h["gtype"], h["cycle"], h["time"], h["multivar"] = take("i"), take("Q"), take("d"), take("i")
    h["keys"] = np.array(take("Q", h["nelems"]) if h["nelems"] > 1 else [take("Q")])
    h["header_bytes"] = off
    h["data"] = np.frombuffer(b[off:], dtype=np.float64)
    return h


def test_gio_layout_and_round_trip(tmp_path):
    g = ga.BoxGrid(2, 4, [3, 2])
    u = np.sin(g.xyz[0]) + g.xyz[1]
    path = tmp_path / "u.000010.out"
    write_state(path, [u], [4, 4], 0,
                g.nelems, time=0.25, cycle=10)
    h = _parse(path)
    assert (h["ivers"], h["dim"], h["nelems"], h["gtype"], h["cycle"], h["time"], h["multivar"]) \
        == (0, 2, 6, 0, 10, 0.25, 0)
    assert np.array_equal(h["porder"], [[4, 4]])
    assert np.array_equal(h["keys"], np.arange(6))
    # GIO::sz_header: (rows*dim + 4) GINT + 2 GSIZET + 1 Ftype + nelems GKEY
    assert h["header_bytes"] == (1 * 2 + 4) * 4 + 2 * 8 + 8 + 6 * 8
    assert np.array_equal(h["data"], u)
    meta, rows, keys, fields = read_state(path)
    assert meta["time"] == 0.25 and meta["cycle"] == 10 and np.array_equal(rows, [[4, 4]])
    assert np.array_equal(fields[0], u)


def test_gio_version1_multivar_and_keys(tmp_path):
    g = ga.BoxGrid(3, 2, [2, 1, 2])
    comps = [g.xyz[0], g.xyz[1] * 2, -g.xyz[2]]
    rows = np.tile([2, 2, 2], (g.nelems, 1))
    keys = np.arange(g.nelems, dtype=np.uint64)[::-1] * 7
    path = tmp_path / "state.out"
    write_state(path, comps, rows, 0, g.nelems, time=1.5, cycle=3, ivers=1, keys=keys)
    h = _parse(path)
    assert h["ivers"] == 1 and h["multivar"] == 1 and np.array_equal(h["porder"], rows)
    assert np.array_equal(h["keys"], keys)
    assert h["header_bytes"] == (g.nelems * 3 + 4) * 4 + 2 * 8 + 8 + g.nelems * 8
    assert np.array_equal(h["data"], np.concatenate(comps))
    meta, r, k, fields = read_state(path, nfields=3)
    assert np.array_equal(k, keys) and np.array_equal(r, rows)
    for a, b in zip(fields, comps):
        assert np.array_equal(a, b)
