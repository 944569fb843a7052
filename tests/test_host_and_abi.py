"""CPU tests of the product library's host side and of the C ABI surface.

* libgeoflow_hip.so loads and exports every function include/*.h declares;
* host set-up (GLL basis, box grid, GGFX maps) equals the oracle's, the maps
  bit-for-bit, and also equals brute-force matching on tiny grids;
* argument errors come back as status codes, not crashes.
No device work is done here.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle
import geoflow_amd as ga
from geoflow_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(gf_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_capi.lib_path())
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares a signature for every entry point
    assert names == set(_capi.SIGNATURES), names ^ set(_capi.SIGNATURES)


def test_version_string():
    assert b"gfx950" in ga.lib().gf_version()


def test_ctx_create_without_device_is_an_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert ga.lib().gf_ctx_create(0, None, ctypes.byref(h)) != 0


@pytest.mark.parametrize("p", range(1, 13))
def test_gll_basis_matches_oracle(p):
    xo, wo, Do = oracle.gll_basis(p)
    xg, wg, Dg = ga.gll_basis(p)
    assert np.array_equal(xo, xg)
    assert np.array_equal(wo, wg)
    assert np.abs(Do - Dg).max() <= 4 * np.finfo(float).eps * np.abs(Do).max()


CASES = [(2, 4, [16, 16], [6] * 4), (2, 6, [4, 4], [0] * 4), (3, 3, [3, 4, 5], [6] * 6),
         (3, 7, [3, 3, 3], [0] * 6), (3, 2, [2, 3, 2], [6, 6, 0, 0, 6, 6]),
         (3, [2, 3, 4], [2, 2, 2], [6] * 6), (3, 4, [2, 2, 2], [6, 6, 6, 6, 0, 0])]


@pytest.mark.parametrize("dim,p,ne,bdy", CASES)
def test_box_grid_and_ggfx_match_oracle(dim, p, ne, bdy):
    og = oracle.Grid.box(dim, p, ne, bdy=bdy)
    bg = ga.BoxGrid(dim, p, ne, bdy=bdy)
    assert bg.ndof == og.ndof
    assert bg.minnodedist == og.minnodedist
    for d, c in enumerate("xyz"[:dim]):
        assert np.array_equal(bg.xyz[d], og.get(c))
        assert np.array_equal(bg.xyzp[d], og.get(c + "p"))
    assert np.array_equal(bg.mask, og.get("mask"))
    nn = int(np.prod(bg.N))
    for d in range(dim):
        assert np.array_equal(np.repeat(bg.edx[d::dim], nn), og.get("dxidx_%d0" % d))
    assert np.array_equal(np.repeat(bg.ejac, nn), og.get("jac"))
    offs, mem = og.ggfx()
    assert np.array_equal(offs, bg.ggfx_offs)
    assert np.array_equal(mem, bg.ggfx_members)


def test_ggfx_bruteforce_tiny():
    bg = ga.BoxGrid(3, 2, [2, 2, 2], bdy=[6, 6, 0, 0, 6, 6])
    X = np.stack(bg.xyzp, 1)
    tol = bg.tol
    for i in range(bg.ndof):
        hit = np.all((X[i] - tol <= X + tol) & (X[i] + tol >= X - tol), axis=1)
        assert np.array_equal(np.flatnonzero(hit), bg.ggfx_members[bg.ggfx_offs[i]:bg.ggfx_offs[i + 1]])


def test_ggfx_generic_coordinates():
    """Scattered points with planted coincidences (not a lattice)."""
    rng = np.random.default_rng(1)
    base = rng.uniform(-3, 3, size=(400, 3))
    pts = np.concatenate([base, base[:150] + 1e-9, base[:40] - 2e-9])
    perm = rng.permutation(len(pts))
    pts = pts[perm]
    offs, mem = ga.ggfx_build([pts[:, 0], pts[:, 1], pts[:, 2]], 1e-7)
    for i in range(len(pts)):
        d = np.abs(pts - pts[i]).max(axis=1)
        assert np.array_equal(np.flatnonzero(d <= 2e-7), mem[offs[i]:offs[i + 1]])


def test_argument_errors_are_status_codes():
    L = ga.lib()
    D = np.zeros(4)
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.gf_gll_basis(0, D.ctypes.data_as(dp), D.ctypes.data_as(dp), D.ctypes.data_as(dp)) == _capi.GF_ERR_ARG
    assert L.gf_gll_basis(3, None, None, None) == _capi.GF_ERR_ARG
    ia = (ctypes.c_int * 3)(2, 2, 2)
    assert L.gf_box_grid(4, ia, ia, None, None, ia, None, None, None, None, None, None) == _capi.GF_ERR_ARG
    # a group larger than max_dup is rejected
    x = np.zeros(5)
    with pytest.raises(_capi.GeoflowError):
        ga.ggfx_build([x, x, x], 1e-6, max_dup=4)
