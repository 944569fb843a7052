"""CPU oracle for the GeoFLOW SEM operator path -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it.  The product library
(``geoflow_amd``) never imports, links or calls anything here.

It wraps ``liboracle.so`` (built from ``gf_oracle.cpp`` by ``oracle/Makefile``),
a C++ restatement of the reference algorithms in the reference's operation
order; see ``gf_oracle.h`` for the per-function citations.

What pins each oracle function (closed forms, the reference's own test
criteria, brute force) is checked in ``tests/test_oracle_pins.py``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

BDY_DIRICHLET = 0
BDY_PERIODIC = 6
BDY_NONE = 7

_dp = ctypes.POINTER(ctypes.c_double)
_szp = ctypes.POINTER(ctypes.c_size_t)


class OrCGTraits(ctypes.Structure):
    _fields_ = [("maxit", ctypes.c_int), ("tol", ctypes.c_double),
                ("normtype", ctypes.c_int), ("precond", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.or_grid_box_create.restype = ctypes.c_void_p
        L.or_grid_box_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                         ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        L.or_grid_icos_create.restype = ctypes.c_void_p
        L.or_grid_icos_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double]
        L.or_grid_icos_create2.restype = ctypes.c_void_p
        L.or_grid_icos_create2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double]
        L.or_grid_terrain_create.restype = ctypes.c_void_p
        L.or_grid_terrain_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double]
        L.or_grid_terrain3_create.restype = ctypes.c_void_p
        L.or_grid_terrain3_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp,
                                              _dp, ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                              ctypes.c_double]
        L.or_grid_set_coords2d.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.or_burgers_rhs_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                             ctypes.POINTER(_dp)]
        L.or_burgers_step_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double,
                                              ctypes.POINTER(_dp), ctypes.POINTER(_dp)]
        L.or_grid_destroy.argtypes = [ctypes.c_void_p]
        for f in ("or_grid_ndof", "or_grid_nelems"):
            getattr(L, f).restype = ctypes.c_size_t
            getattr(L, f).argtypes = [ctypes.c_void_p]
        for f in ("or_grid_dim", "or_grid_gtype"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [ctypes.c_void_p]
        for f in ("or_grid_minnodedist", "or_grid_volume"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.or_grid_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
        L.or_grid_ggfx.argtypes = [ctypes.c_void_p, _szp, _szp, _szp]
        L.or_gll_basis.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_dmxm.argtypes = [_dp, _dp, ctypes.c_size_t, ctypes.c_size_t, _dp,
                              ctypes.c_size_t, ctypes.c_size_t]
        L.or_grefderiv_raw.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_size_t, ctypes.POINTER(_dp), ctypes.c_int,
                                       ctypes.c_int, _dp, _dp]
        L.or_grefderiv.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.or_grefdiv.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp), ctypes.c_int, _dp]
        L.or_deriv.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.or_integrate.restype = ctypes.c_double
        L.or_integrate.argtypes = [ctypes.c_void_p, _dp]
        L.or_mass_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_imass_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_dealias.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.or_helmholtz_apply.argtypes = [ctypes.c_void_p, _dp, ctypes.c_size_t, _dp, _dp, _dp]
        L.or_dss.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.or_imult.argtypes = [ctypes.c_void_p, _dp]
        L.or_advect.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.POINTER(_dp), _dp]
        L.or_constrain2sphere.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.or_burgers_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]
        L.or_burgers_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.POINTER(_dp)]
        L.or_cg_solve.restype = ctypes.c_int
        L.or_cg_solve.argtypes = [ctypes.c_void_p, _dp, ctypes.c_size_t, _dp, _dp, _dp, _dp,
                                  ctypes.POINTER(OrCGTraits), ctypes.POINTER(ctypes.c_int), _dp]
        L.or_helmholtz_diag_inv.argtypes = [ctypes.c_void_p, _dp, ctypes.c_size_t, _dp, _dp]
        L.or_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _pp(arrs):
    arr = (_dp * len(arrs))()
    for i, a in enumerate(arrs):
        arr[i] = _p(a) if a is not None else None
    return arr


def set_threads(n):
    lib().or_set_threads(int(n))


def gll_basis(p):
    N = p + 1
    xi = np.zeros(N)
    w = np.zeros(N)
    D = np.zeros(N * N)
    if lib().or_gll_basis(p, _p(xi), _p(w), _p(D)) != 0:
        raise ValueError("bad order")
    return xi, w, D.reshape(N, N, order="F")


def dmxm(A, B):
    """C = A @ B with dmxm's summation order (A, B as 2-D numpy arrays)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    C = np.zeros((A.shape[0], B.shape[1]), order="F")
    lib().or_dmxm(C.ctypes.data_as(_dp), A.ctypes.data_as(_dp), A.shape[0], A.shape[1],
                  B.ctypes.data_as(_dp), B.shape[0], B.shape[1])
    return C


def grefderiv_raw(dim, N, ne, Dmats, idir, dotrans, u):
    Ns = (ctypes.c_int * 3)(*(list(N) + [1] * (3 - len(N))))
    Dc = [np.ascontiguousarray(np.asarray(D, dtype=np.float64).reshape(-1, order="F"))
          for D in Dmats]
    while len(Dc) < 3:
        Dc.append(np.zeros(1))
    du = np.zeros_like(u)
    lib().or_grefderiv_raw(dim, Ns, ne, _pp(Dc), idir, int(dotrans), _p(u), _p(du))
    return du


class Grid:
    """An oracle grid (box or icosahedral)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle grid construction failed")
        self.h = ctypes.c_void_p(handle)
        L = lib()
        self.ndof = L.or_grid_ndof(self.h)
        self.nelems = L.or_grid_nelems(self.h)
        self.dim = L.or_grid_dim(self.h)
        self.gtype = L.or_grid_gtype(self.h)
        self.minnodedist = L.or_grid_minnodedist(self.h)
        self.volume = L.or_grid_volume(self.h)

    @classmethod
    def box(cls, dim, p, ne, P0=None, P1=None, bdy=None, tol_fac=0.001):
        if np.isscalar(p):
            p = [p] * dim
        P0 = np.asarray(P0 if P0 is not None else [0.0] * 3, dtype=np.float64)
        P1 = np.asarray(P1 if P1 is not None else [1.0] * 3, dtype=np.float64)
        P0 = np.ascontiguousarray(np.pad(P0, (0, 3 - len(P0))))
        P1 = np.ascontiguousarray(np.pad(P1, (0, 3 - len(P1))))
        if bdy is None:
            bdy = [BDY_PERIODIC] * (2 * dim)
        pa = (ctypes.c_int * 3)(*(list(p) + [0] * (3 - len(p))))
        na = (ctypes.c_int * 3)(*(list(ne) + [1] * (3 - len(ne))))
        ba = (ctypes.c_int * 6)(*(list(bdy) + [BDY_NONE] * (6 - len(bdy))))
        h = lib().or_grid_box_create(dim, pa, na, _p(P0), _p(P1), ba, tol_fac)
        g = cls(h)
        g.p = list(p)
        g.ne = list(ne)
        g.P0, g.P1, g.bdy = P0, P1, list(bdy)
        return g

    @classmethod
    def icos(cls, p, ilevel, radius=1.0, tol_fac=0.001, refine=0):
        """refine 0: GICOS_BISECTION, 1: GICOS_LAGRANGIAN."""
        g = cls(lib().or_grid_icos_create2(p, ilevel, refine, radius, tol_fac))
        g.p = [p, p]
        return g

    @classmethod
    def terrain(cls, p, ne, amp, P0=(0.0, 0.0), P1=(1.0, 1.0), bdy=None, ref_inverse=True,
                tol_fac=0.001):
        """2-D box of GE_DEFORMED quads under the terrain-following map
        y' = y + h(x)(P1y - y)/H, h = amp (1 - cos(2 pi (x-P0x)/Lx))/2;
        ref_inverse keeps the reference's transposed 2x2 dXi/dX."""
        if bdy is None:
            bdy = [BDY_PERIODIC, BDY_PERIODIC, 0, 0]
        P0 = np.ascontiguousarray(P0, dtype=np.float64)
        P1 = np.ascontiguousarray(P1, dtype=np.float64)
        na = (ctypes.c_int * 2)(*ne)
        ba = (ctypes.c_int * 4)(*bdy)
        g = cls(lib().or_grid_terrain_create(p, na, _p(P0), _p(P1), ba, amp, int(bool(ref_inverse)),
                                             tol_fac))
        g.p = [p, p]
        g.ne = list(ne)
        return g

    @classmethod
    def terrain3(cls, p, ne, amp, P0=(0.0, 0.0, 0.0), P1=(1.0, 1.0, 1.0), bdy=None,
                 tol_fac=0.001):
        """3-D box of GE_DEFORMED hexahedra, bottom raised to
        h = amp (1 - cos(2 pi x/Lx))(1 - cos(2 pi y/Ly))/4."""
        if bdy is None:
            bdy = [BDY_PERIODIC] * 4 + [0, 0]
        P0 = np.ascontiguousarray(P0, dtype=np.float64)
        P1 = np.ascontiguousarray(P1, dtype=np.float64)
        g = cls(lib().or_grid_terrain3_create(p, (ctypes.c_int * 3)(*ne), _p(P0), _p(P1),
                                              (ctypes.c_int * 6)(*bdy), amp, tol_fac))
        g.p = [p] * 3
        g.ne = list(ne)
        return g

    def set_coords(self, x, y, ref_inverse=True, tol_fac=0.001):
        """New node coordinates for a 2-D deformed grid without periodic
        directions (the end of GGrid::add_terrain): geometry, GGFX, mass and
        volume recomputed."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if lib().or_grid_set_coords2d(self.h, _p(x), _p(y), int(bool(ref_inverse)), tol_fac) != 0:
            raise ValueError("set_coords: 2-D deformed grids only")
        self.minnodedist = lib().or_grid_minnodedist(self.h)
        self.volume = lib().or_grid_volume(self.h)
        return self

    def __del__(self):
        try:
            lib().or_grid_destroy(self.h)
        except Exception:
            pass

    def get(self, which):
        out = np.zeros(self.ndof)
        if lib().or_grid_get(self.h, which.encode(), _p(out)) != 0:
            raise KeyError(which)
        return out

    def ggfx(self):
        nm = ctypes.c_size_t(0)
        lib().or_grid_ggfx(self.h, None, None, ctypes.byref(nm))
        offs = np.zeros(self.ndof + 1, dtype=np.uint64)
        mem = np.zeros(nm.value, dtype=np.uint64)
        lib().or_grid_ggfx(self.h, offs.ctypes.data_as(_szp), mem.ctypes.data_as(_szp),
                           ctypes.byref(nm))
        return offs.astype(np.int64), mem.astype(np.int64)

    # -- operators ---------------------------------------------------------
    def grefderiv(self, u, idir, dotrans=False):
        du = np.zeros(self.ndof)
        lib().or_grefderiv(self.h, _p(u), idir, int(dotrans), _p(du))
        return du

    def grefdiv(self, us, dotrans=True):
        out = np.zeros(self.ndof)
        lib().or_grefdiv(self.h, _pp(list(us) + [None] * (3 - len(us))), int(dotrans), _p(out))
        return out

    def deriv(self, u, idir):
        du = np.zeros(self.ndof)
        lib().or_deriv(self.h, _p(u), idir, _p(du))
        return du

    def integrate(self, u):
        return lib().or_integrate(self.h, _p(np.ascontiguousarray(u, dtype=np.float64)))

    def mass_apply(self, u):
        o = np.zeros(self.ndof)
        lib().or_mass_apply(self.h, _p(u), _p(o))
        return o

    def imass_apply(self, u):
        o = np.zeros(self.ndof)
        lib().or_imass_apply(self.h, _p(u), _p(o))
        return o

    def dealias(self, v1, v2, pq):
        """GGrid::dealias: v1 v2 through the order-pq GLL grid."""
        o = np.zeros(self.ndof)
        lib().or_dealias(self.h, int(pq), _p(np.ascontiguousarray(v1)),
                         _p(np.ascontiguousarray(v2)), _p(o))
        return o

    def helmholtz(self, u, p=1.0, q=None):
        pa = np.atleast_1d(np.asarray(p, dtype=np.float64))
        o = np.zeros(self.ndof)
        lib().or_helmholtz_apply(self.h, _p(pa), pa.size, _p(q), _p(u), _p(o))
        return o

    def dss(self, u, op="sum"):
        ops = {"sum": 0, "smooth": 1, "max": 2, "min": 3}
        v = np.array(u, dtype=np.float64, copy=True)
        lib().or_dss(self.h, ops[op], _p(v))
        return v

    def imult(self):
        o = np.zeros(self.ndof)
        lib().or_imult(self.h, _p(o))
        return o

    def advect(self, p, us):
        o = np.zeros(self.ndof)
        lib().or_advect(self.h, _p(p), len(us), _pp(us), _p(o))
        return o

    def constrain2sphere(self, vs):
        """In-place GMTK::constrain2sphere of the three components vs."""
        lib().or_constrain2sphere(self.h, _p(vs[0]), _p(vs[1]), _p(vs[2]))
        return vs

    def burgers_rhs(self, us, nu, mode=0):
        # heat (2) and nonlinear (0) modes produce one output per state component
        outs = [np.zeros(self.ndof) for _ in range(1 if mode == 1 else len(us))]
        lib().or_burgers_rhs(self.h, mode, nu, len(us), _pp(us), _pp(outs))
        return outs

    def burgers_step(self, us, nu, dt, t=0.0, mode=0, order=4):
        """Advance the state list in place by one ExRK step."""
        lib().or_burgers_step(self.h, mode, nu, order, len(us), t, dt, _pp(us))
        return us

    def burgers_rhs_adjoint(self, us, ws, lams, nu):
        """lams[k] += (dF/du)^T w at the state us (nonlinear mode), in place."""
        lib().or_burgers_rhs_adjoint(self.h, nu, len(us), _pp(us), _pp(ws), _pp(lams))
        return lams

    def burgers_step_adjoint(self, un, lams, nu, dt, order=4):
        """lams (dJ/du_{n+1}) becomes dJ/du_n for one step from un, in place."""
        lib().or_burgers_step_adjoint(self.h, nu, order, len(un), dt, _pp(un), _pp(lams))
        return lams

    def helmholtz_diag_inv(self, p=1.0, q=None):
        pa = np.atleast_1d(np.asarray(p, dtype=np.float64))
        o = np.zeros(self.ndof)
        lib().or_helmholtz_diag_inv(self.h, _p(pa), pa.size, _p(q), _p(o))
        return o

    def cg_solve(self, b, x0=None, xb=None, p=1.0, q=None, maxit=512, tol=1e-6,
                 normtype=0, precond=0):
        pa = np.atleast_1d(np.asarray(p, dtype=np.float64))
        x = np.zeros(self.ndof) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        tr = OrCGTraits(maxit, tol, normtype, precond)
        it = ctypes.c_int(0)
        res = np.zeros(max(maxit, 1))
        ret = lib().or_cg_solve(self.h, _p(pa), pa.size, _p(q), _p(b), _p(xb), _p(x),
                                ctypes.byref(tr), ctypes.byref(it), _p(res))
        return x, ret, it.value, res[: it.value]
