"""Host-side grid construction through the C ABI (no device work).

``BoxGrid`` mirrors GGridBox (src/cdg/grid/ggrid_box.ipp) plus the GGFX set-up
of ``init_ggfx`` (src/apps/geoflow_cdg.cpp:731-774): node coordinates,
regular-element metrics, boundary mask, periodized coordinates and the
coincidence groups matched with tolerance 0.001 * minnodedist.

``IcosGrid`` mirrors the 2-D GGridIcos (src/cdg/grid/ggrid_icos.ipp): quads on
the sphere with per-node embedded metrics, and the same GGFX set-up.

``TerrainBoxGrid`` is a 2-D box of GE_DEFORMED quads whose bottom follows a
terrain profile (the result of GGrid::add_terrain, src/cdg/grid/ggrid.ipp:1486-1548,
with an analytic terrain-following map), with per-node deformed metrics.
"""
import ctypes

import numpy as np

from ._capi import (GF_BDY_PERIODIC, GF_BDY_DIRICHLET, GF_GE_REGULAR, GF_GE_2DEMBEDDED,
                    GF_GE_DEFORMED, check, lib)

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def gll_basis(p):
    """GLL nodes, weights and derivative matrix D (N x N, D[l, j] = dphi_j/dxi(xi_l))."""
    N = p + 1
    xi, w, D = np.zeros(N), np.zeros(N), np.zeros(N * N)
    check(lib().gf_gll_basis(p, _p(xi), _p(w), _p(D)))
    return xi, w, D.reshape(N, N, order="F")


def box_grid(dim, p, ne, P0, P1, bdy):
    """Run gf_box_grid; returns dict of host arrays."""
    p = [p] * dim if np.isscalar(p) else list(p)
    ne = list(ne)
    nn = int(np.prod([q + 1 for q in p]))
    nel = int(np.prod(ne))
    ndof = nn * nel
    xyz = [np.zeros(ndof) for _ in range(dim)]
    xyzp = [np.zeros(ndof) for _ in range(dim)]
    mask = np.zeros(ndof)
    edx = np.zeros(nel * dim)
    ejac = np.zeros(nel)
    md = ctypes.c_double(0)
    ia = lambda v: (ctypes.c_int * len(v))(*v)
    P0a = np.ascontiguousarray(P0[:dim], dtype=np.float64)
    P1a = np.ascontiguousarray(P1[:dim], dtype=np.float64)
    xp = (_dp * dim)(*[_p(a) for a in xyz])
    xpp = (_dp * dim)(*[_p(a) for a in xyzp])
    check(lib().gf_box_grid(dim, ia(p), ia(ne), _p(P0a), _p(P1a), ia(list(bdy)), xp, xpp,
                            _p(mask), _p(edx), _p(ejac), ctypes.byref(md)))
    return dict(xyz=xyz, xyzp=xyzp, mask=mask, edx=edx, ejac=ejac, minnodedist=md.value,
                ndof=ndof, nelems=nel, p=p, ne=ne)


def ggfx_build(coords, tol, max_dup=0):
    """GGFX coincidence groups: (offs[ndof+1], members) as int64 arrays."""
    coords = [np.ascontiguousarray(c, dtype=np.float64) for c in coords]
    ndof = coords[0].size
    cp = (_dp * len(coords))(*[_p(c) for c in coords])
    nm = ctypes.c_size_t(0)
    offs = np.zeros(ndof + 1, dtype=np.int64)
    check(lib().gf_ggfx_build(ndof, len(coords), cp, tol, max_dup, offs.ctypes.data_as(_i64p),
                              None, ctypes.byref(nm)))
    mem = np.zeros(max(nm.value, 1), dtype=np.int64)
    check(lib().gf_ggfx_build(ndof, len(coords), cp, tol, max_dup, offs.ctypes.data_as(_i64p),
                              mem.ctypes.data_as(_i64p), ctypes.byref(nm)))
    return offs, mem[: nm.value]


def icos_grid(p, ilevel, refine=0, radius=1.0):
    """Run gf_icos_grid; returns dict of host arrays."""
    ne = ctypes.c_size_t(0)
    check(lib().gf_icos_grid(p, ilevel, refine, radius, ctypes.byref(ne), None, None, None,
                             None))
    nel = ne.value
    ndof = nel * (p + 1) ** 2
    xyz = [np.zeros(ndof) for _ in range(3)]
    dxidx = [[np.zeros(ndof) for _ in range(3)] for _ in range(2)]
    jac = np.zeros(ndof)
    md = ctypes.c_double(0)
    xp = (_dp * 3)(*[_p(a) for a in xyz])
    dp = (_dp * 6)(*[_p(dxidx[i][j]) for i in range(2) for j in range(3)])
    check(lib().gf_icos_grid(p, ilevel, refine, radius, ctypes.byref(ne), xp, dp, _p(jac),
                             ctypes.byref(md)))
    return dict(xyz=xyz, node_dxidx=dxidx, node_jac=jac, minnodedist=md.value, ndof=ndof,
                nelems=nel)


class IcosGrid:
    """GGridIcos (2-D) equivalent: quads on the sphere of the given radius,
    GE_2DEMBEDDED metrics per node; refine 0 is GICOS_BISECTION (ilevel
    bisections), 1 GICOS_LAGRANGIAN (ilevel rows per base triangle)."""

    gtype = GF_GE_2DEMBEDDED

    def __init__(self, p, ilevel, radius=1.0, refine=0, tol_fac=0.001):
        self.dim = 2
        g = icos_grid(p, ilevel, refine, radius)
        self.__dict__.update(g)
        self.p = [p, p]
        self.N = [p + 1, p + 1]
        self.ilevel, self.refine, self.radius = ilevel, refine, radius
        self.basis = [gll_basis(p)] * 2
        self.mask = None
        self.tol = tol_fac * self.minnodedist
        # GGridIcos::max_duplicates: 6 quads meet at a base-icosahedron vertex
        self.ggfx_offs, self.ggfx_members = ggfx_build(self.xyz, self.tol, max_dup=6)


def terrain_box_grid(p, ne, P0, P1, bdy, amp, reference_inverse=True):
    """Run gf_terrain_box_grid; returns dict of host arrays (node_dxidx[i][j],
    i, j < 2)."""
    ne = [int(q) for q in ne]
    nel = ne[0] * ne[1]
    ndof = nel * (p + 1) ** 2
    xyz = [np.zeros(ndof) for _ in range(2)]
    xyzp = [np.zeros(ndof) for _ in range(2)]
    mask = np.zeros(ndof)
    dxidx = [[np.zeros(ndof) for _ in range(2)] for _ in range(2)]
    jac = np.zeros(ndof)
    md = ctypes.c_double(0)
    ia = lambda v: (ctypes.c_int * len(v))(*v)
    P0a = np.ascontiguousarray(P0[:2], dtype=np.float64)
    P1a = np.ascontiguousarray(P1[:2], dtype=np.float64)
    xp = (_dp * 2)(*[_p(a) for a in xyz])
    xpp = (_dp * 2)(*[_p(a) for a in xyzp])
    dp = (_dp * 4)(*[_p(dxidx[i][j]) for i in range(2) for j in range(2)])
    check(lib().gf_terrain_box_grid(p, ia(ne), _p(P0a), _p(P1a), ia(list(bdy)), float(amp),
                                    int(bool(reference_inverse)), xp, xpp, _p(mask), dp, _p(jac),
                                    ctypes.byref(md)))
    return dict(xyz=xyz, xyzp=xyzp, mask=mask, node_dxidx=dxidx, node_jac=jac,
                minnodedist=md.value, ndof=ndof, nelems=nel)


def deformed_geometry2d(p, xyz, reference_inverse=True):
    """Run gf_deformed_geometry2d: per-node dXi/dX (node_dxidx[i][j]) and
    Jacobian of 2-D elements from their node coordinates."""
    xyz = [np.ascontiguousarray(a, dtype=np.float64) for a in xyz[:2]]
    ndof = xyz[0].size
    nel = ndof // (p + 1) ** 2
    dxidx = [[np.zeros(ndof) for _ in range(2)] for _ in range(2)]
    jac = np.zeros(ndof)
    xp = (_dp * 2)(*[_p(a) for a in xyz])
    dp = (_dp * 4)(*[_p(dxidx[i][j]) for i in range(2) for j in range(2)])
    check(lib().gf_deformed_geometry2d(p, nel, xp, int(bool(reference_inverse)), dp, _p(jac)))
    return dxidx, jac


class TerrainBoxGrid:
    """2-D box [P0, P1] of GE_DEFORMED quads with the bottom raised to
    h(x) = amp (1 - cos(2 pi (x - P0x)/Lx)) / 2 and every node moved by the
    terrain-following map y' = y + h(x) (P1y - y)/(P1y - P0y).  With
    reference_inverse (default) the metrics are the reference's, whose 2x2
    inverse is transposed (GElem_base::inv, gelem_base.cpp:1249-1266);
    otherwise dXi/dX is the true inverse of dX/dXi."""

    gtype = GF_GE_DEFORMED

    def __init__(self, p, ne, amp=0.1, P0=(0.0, 0.0), P1=(1.0, 1.0), bdy=None,
                 reference_inverse=True, tol_fac=0.001):
        if bdy is None:
            bdy = [GF_BDY_PERIODIC, GF_BDY_PERIODIC, GF_BDY_DIRICHLET, GF_BDY_DIRICHLET]
        self.dim = 2
        g = terrain_box_grid(p, ne, np.asarray(P0, dtype=np.float64),
                             np.asarray(P1, dtype=np.float64), bdy, amp, reference_inverse)
        self.__dict__.update(g)
        self.p = [p, p]
        self.N = [p + 1, p + 1]
        self.ne = [int(q) for q in ne]
        self.bdy = list(bdy)
        self.amp = amp
        self.reference_inverse = reference_inverse
        self.P0 = np.asarray(P0, dtype=np.float64)[:2]
        self.P1 = np.asarray(P1, dtype=np.float64)[:2]
        self.basis = [gll_basis(p)] * 2
        self.tol = tol_fac * self.minnodedist
        self.ggfx_offs, self.ggfx_members = ggfx_build(self.xyzp, self.tol, max_dup=4)


class TerrainBoxGrid3D:
    """3-D box [P0, P1] of GE_DEFORMED hexahedra with the bottom raised to
    h(x,y) = amp (1 - cos(2 pi x/Lx)) (1 - cos(2 pi y/Ly)) / 4 and every node
    moved by z' = z + h (P1z - z)/(P1z - P0z); per-node metrics of
    GElem_base::dogeom3d (3x3 inverse, the true one)."""

    gtype = GF_GE_DEFORMED

    def __init__(self, p, ne, amp=0.1, P0=(0.0, 0.0, 0.0), P1=(1.0, 1.0, 1.0), bdy=None,
                 tol_fac=0.001):
        if bdy is None:
            bdy = [GF_BDY_PERIODIC] * 4 + [GF_BDY_DIRICHLET] * 2
        ne = [int(q) for q in ne]
        nel = ne[0] * ne[1] * ne[2]
        ndof = nel * (p + 1) ** 3
        xyz = [np.zeros(ndof) for _ in range(3)]
        xyzp = [np.zeros(ndof) for _ in range(3)]
        mask = np.zeros(ndof)
        dxidx = [[np.zeros(ndof) for _ in range(3)] for _ in range(3)]
        jac = np.zeros(ndof)
        md = ctypes.c_double(0)
        ia = lambda v: (ctypes.c_int * len(v))(*v)
        P0a = np.ascontiguousarray(P0[:3], dtype=np.float64)
        P1a = np.ascontiguousarray(P1[:3], dtype=np.float64)
        check(lib().gf_terrain_box_grid3d(
            p, ia(ne), _p(P0a), _p(P1a), ia(list(bdy)), float(amp),
            (_dp * 3)(*[_p(a) for a in xyz]), (_dp * 3)(*[_p(a) for a in xyzp]), _p(mask),
            (_dp * 9)(*[_p(dxidx[i][j]) for i in range(3) for j in range(3)]), _p(jac),
            ctypes.byref(md)))
        self.dim = 3
        self.xyz, self.xyzp, self.mask = xyz, xyzp, mask
        self.node_dxidx, self.node_jac = dxidx, jac
        self.minnodedist = md.value
        self.ndof, self.nelems = ndof, nel
        self.p, self.N, self.ne = [p] * 3, [p + 1] * 3, ne
        self.bdy, self.amp = list(bdy), amp
        self.P0, self.P1 = P0a, P1a
        self.basis = [gll_basis(p)] * 3
        self.tol = tol_fac * self.minnodedist
        self.ggfx_offs, self.ggfx_members = ggfx_build(self.xyzp, self.tol, max_dup=8)


class DeformedGrid2D:
    """2-D GE_DEFORMED grid from explicit node coordinates on the elements of
    a box grid (what GGrid::add_terrain leaves, src/cdg/grid/ggrid.ipp:1522-1547):
    geometry recomputed from the coordinates, mask kept from the base grid,
    GGFX matched on the new coordinates (no periodic directions)."""

    gtype = GF_GE_DEFORMED

    def __init__(self, base, xyz, reference_inverse=True, tol_fac=0.001):
        self.dim = 2
        self.p, self.N, self.ne = list(base.p), list(base.N), list(base.ne)
        self.basis = base.basis
        self.nelems, self.ndof = base.nelems, base.ndof
        self.mask = base.mask
        self.reference_inverse = reference_inverse
        self.xyz = [np.ascontiguousarray(a, dtype=np.float64) for a in xyz[:2]]
        self.xyzp = self.xyz
        self.node_dxidx, self.node_jac = deformed_geometry2d(self.p[0], self.xyz,
                                                             reference_inverse)
        d = [np.hypot(np.diff(self.xyz[0]), np.diff(self.xyz[1]))]
        tiny = 1000.0 * np.finfo(np.float64).eps
        self.minnodedist = float(d[0][d[0] > tiny].min())  # GGrid::find_min_dist
        self.tol = tol_fac * self.minnodedist
        self.ggfx_offs, self.ggfx_members = ggfx_build(self.xyzp, self.tol, max_dup=4)


class BoxGrid:
    """GGridBox equivalent: host arrays for a box of regular elements."""

    gtype = GF_GE_REGULAR

    def __init__(self, dim, p, ne, P0=(0.0, 0.0, 0.0), P1=(1.0, 1.0, 1.0), bdy=None,
                 tol_fac=0.001, device_maps=False):
        """device_maps: leave the GGFX maps to the device (gf_box_ggfx_device
        at mesh creation; 3-D boxes periodic in every direction or in none),
        so no host matching runs; ggfx_offs/ggfx_members are then None."""
        if bdy is None:
            bdy = [GF_BDY_PERIODIC] * (2 * dim)
        self.dim = dim
        g = box_grid(dim, p, ne, np.asarray(P0, dtype=np.float64), np.asarray(P1, dtype=np.float64),
                     bdy)
        self.__dict__.update(g)
        self.bdy = list(bdy)
        self.ne = [int(q) for q in ne][:dim]
        self.periodic = [int(bdy[2 * d] == GF_BDY_PERIODIC and bdy[2 * d + 1] == GF_BDY_PERIODIC)
                         for d in range(dim)]
        self.P0 = np.asarray(P0, dtype=np.float64)[:dim]
        self.P1 = np.asarray(P1, dtype=np.float64)[:dim]
        self.N = [q + 1 for q in self.p]
        self.basis = [gll_basis(q) for q in self.p]
        self.tol = tol_fac * self.minnodedist
        self.device_maps = bool(device_maps)
        if self.device_maps:
            if dim != 3 or len(set(self.p)) != 1 or sum(self.periodic) not in (0, 3):
                raise ValueError("device GGFX maps: 3-D isotropic boxes, periodic in every "
                                 "direction or in none")
            self.ggfx_offs = self.ggfx_members = None
        else:
            # GGFX::init uses 3 coordinates; missing ones are zero (geoflow_cdg.cpp:745-753)
            self.ggfx_offs, self.ggfx_members = ggfx_build(self.xyzp, self.tol,
                                                           max_dup=2 ** dim)
