"""Device-side mirror of the reference operator API over the C ABI.

Fields are torch float64 CUDA tensors holding element-contiguous local
vectors; every operator call goes straight to libgeoflow_hip.so on the
context's stream (torch's current stream).  There is no CPU path.
"""
import ctypes

import numpy as np
import torch

from . import _capi
from ._capi import (CGParams, CGResult, MeshDesc, GF_GE_REGULAR, GF_GE_2DEMBEDDED, GF_GE_DEFORMED,
                    check, lib)

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _ptrs(ts):
    return (_vp * len(ts))(*[(t.data_ptr() if t is not None else None) for t in ts])


def _hp(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Context:
    """gf_ctx bound to a device and torch's current stream on it."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("geoflow_amd device operators need a GPU")
        self.device = device
        torch.cuda.set_device(device)
        stream = torch.cuda.current_stream(device).cuda_stream
        h = _vp()
        check(lib().gf_ctx_create(device, ctypes.c_void_p(stream), ctypes.byref(h)))
        self.h = h

    def sync(self):
        check(lib().gf_ctx_synchronize(self.h), self.h)

    def peak(self, which):
        """Measured TFLOP/s of "mfma_f32", "mfma_f64" or "valu_f64"."""
        out = ctypes.c_double()
        check(lib().gf_peak_probe(self.h, which.encode(), ctypes.byref(out)), self.h)
        return out.value

    def __del__(self):
        try:
            lib().gf_ctx_destroy(self.h)
        except Exception:
            pass


def _q_args(q):
    """Mass-term factor as (field, scalar): a tensor is a per-node field, a
    number a scalar, None no mass term."""
    if q is None:
        return None, 0.0
    if torch.is_tensor(q):
        return q, 0.0
    return None, float(q)


def box_ggfx_device(ctx, N, ne, periodic):
    """gf_box_ggfx_device: the GGFX CSR maps of a 3-D box built on the
    device, returned as host int64 arrays (offs, members)."""
    nea = (ctypes.c_int * 3)(*[int(q) for q in ne])
    pa = (ctypes.c_int * 3)(*[int(q) for q in periodic])
    ndof = int(np.prod(ne)) * N ** 3
    offs = torch.empty(ndof + 1, dtype=torch.int64, device="cuda")
    nm = ctypes.c_size_t(0)
    check(lib().gf_box_ggfx_device(ctx.h, int(N), nea, pa, _ptr(offs), None, ctypes.byref(nm)),
          ctx.h)
    mem = torch.empty(max(nm.value, 1), dtype=torch.int64, device="cuda")
    check(lib().gf_box_ggfx_device(ctx.h, int(N), nea, pa, _ptr(offs), _ptr(mem),
                                   ctypes.byref(nm)), ctx.h)
    torch.cuda.synchronize()
    return offs.cpu().numpy(), mem[:nm.value].cpu().numpy()


class Mesh:
    """Device-resident mesh: GLL basis, metrics, mask and GGFX groups.

    Built from a host grid object exposing dim, N, nelems, basis, and either
    edx/ejac (regular) or node metrics (embedded), plus ggfx_offs/members.
    """

    def __init__(self, ctx, grid, gtype=None, node_dxidx=None, node_jac=None,
                 xyz=None, use_mask=True, structured=True):
        self.ctx = ctx
        self.grid = grid
        self.dim = grid.dim
        if gtype is None:
            gtype = getattr(grid, "gtype", GF_GE_REGULAR)
        if gtype in (GF_GE_2DEMBEDDED, GF_GE_DEFORMED):
            node_dxidx = grid.node_dxidx if node_dxidx is None else node_dxidx
            node_jac = grid.node_jac if node_jac is None else node_jac
        self.gtype = gtype
        d = MeshDesc()
        d.dim = grid.dim
        d.gtype = gtype
        for k in range(3):
            d.N[k] = grid.N[k] if k < grid.dim else 1
        d.nelems = grid.nelems
        keep = []

        def host(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_dp)

        for k in range(grid.dim):
            xi, w, D = grid.basis[k]
            d.D[k] = host(D.reshape(-1, order="F"))
            d.W[k] = host(w)
        if gtype == GF_GE_REGULAR:
            d.elem_dxidx = host(grid.edx)
            d.elem_jac = host(grid.ejac)
        else:
            for i in range(len(node_dxidx)):
                for j in range(len(node_dxidx[i])):
                    d.node_dxidx[i * 3 + j] = host(node_dxidx[i][j])
            d.node_jac = host(node_jac)
        coords = xyz if xyz is not None else getattr(grid, "xyz", None)
        if coords is not None:
            for k in range(len(coords)):
                d.xyz[k] = host(coords[k])
        if use_mask and getattr(grid, "mask", None) is not None:
            d.mask = host(grid.mask)
        devmaps = getattr(grid, "device_maps", False)
        if devmaps:
            d.box_device_maps = 1
            bl = getattr(grid, "box_list", None)
            if bl is not None:
                arr = (ctypes.c_int * len(bl))(*[int(v) for v in bl])
                keep.append(arr)
                d.box_count = len(bl) // 6
                d.box_list = arr
        else:
            offs = np.ascontiguousarray(grid.ggfx_offs, dtype=np.int64)
            mem = np.ascontiguousarray(grid.ggfx_members, dtype=np.int64)
            keep += [offs, mem]
            d.ggfx_offs = offs.ctypes.data_as(_i64p)
            d.ggfx_members = mem.ctypes.data_as(_i64p)
        if ((structured or devmaps) and gtype == GF_GE_REGULAR and grid.dim == 3
                and getattr(grid, "ne", None) is not None):
            for k in range(3):
                d.box_ne[k] = grid.ne[k]
                d.box_periodic[k] = grid.periodic[k]
        h = _vp()
        check(lib().gf_mesh_create(ctx.h, ctypes.byref(d), ctypes.byref(h)), ctx.h)
        self.h = h
        self.ndof = lib().gf_mesh_ndof(h)
        self.flags = lib().gf_mesh_flags(h)

    def __del__(self):
        try:
            lib().gf_mesh_destroy(self.h)
        except Exception:
            pass

    # -- helpers -----------------------------------------------------------
    def _c(self, rc):
        return check(rc, self.ctx.h)

    def new(self, dtype=torch.float64):
        return torch.zeros(self.ndof, dtype=dtype, device="cuda")

    def get(self, which):
        out = np.zeros(self.ndof)
        self._c(lib().gf_mesh_get(self.h, which.encode(), _hp(out)))
        return out

    # -- operators (mirror GGrid / GMTK / GMass / GHelmholtz / GGFX) -------
    def grefderiv(self, u, idir, dotrans=False, out=None):
        out = self.new() if out is None else out
        self._c(lib().gf_grefderiv(self.h, idir, int(dotrans), _ptr(u), _ptr(out)))
        return out

    def grefderivs(self, u, dotrans=False):
        outs = [self.new() for _ in range(self.dim)]
        self._c(lib().gf_grefderivs(self.h, int(dotrans), _ptr(u), _ptrs(outs)))
        return outs

    def grefdiv(self, us, dotrans=True, out=None):
        out = self.new() if out is None else out
        us = list(us) + [None] * (3 - len(us))
        self._c(lib().gf_grefdiv(self.h, int(dotrans), _ptrs(us), _ptr(out)))
        return out

    def deriv(self, u, idir, out=None):
        out = self.new() if out is None else out
        self._c(lib().gf_deriv(self.h, idir, _ptr(u), _ptr(out)))
        return out

    def grad(self, u):
        # embedded 2-D meshes return all three Cartesian derivatives
        ncomp = 3 if self.gtype == GF_GE_2DEMBEDDED else self.dim
        outs = [self.new() for _ in range(ncomp)]
        self._c(lib().gf_grad(self.h, _ptr(u), _ptrs(outs)))
        return outs

    def mass_apply(self, u, out=None):
        out = self.new() if out is None else out
        self._c(lib().gf_mass_apply(self.h, _ptr(u), _ptr(out)))
        return out

    def imass_apply(self, u, out=None):
        out = self.new() if out is None else out
        self._c(lib().gf_imass_apply(self.h, _ptr(u), _ptr(out)))
        return out

    def integrate(self, u):
        r = ctypes.c_double(0)
        self._c(lib().gf_integrate(self.h, _ptr(u), ctypes.byref(r)))
        return r.value

    def helmholtz(self, u, p=1.0, q=None, out=None):
        """uo = p L u + q M u; p and q may be scalars or per-node fields, q None
        for the pure Laplacian."""
        out = self.new() if out is None else out
        pf = p if torch.is_tensor(p) else None
        ps = 1.0 if pf is not None else float(p)
        qf, qs = _q_args(q)
        self._c(lib().gf_helmholtz_apply(self.h, ps, _ptr(pf), qs, _ptr(qf), _ptr(u), _ptr(out)))
        return out

    def helmholtz_f32(self, u, p=1.0, q=None, out=None):
        out = self.new(torch.float32) if out is None else out
        qf, qs = _q_args(q)
        self._c(lib().gf_helmholtz_apply_f32(self.h, float(p), qs, _ptr(qf), _ptr(u), _ptr(out)))
        return out

    def dealias(self, v1, v2, pq, out=None):
        """GGrid::dealias: v1 v2 through the order-pq GLL grid."""
        out = self.new() if out is None else out
        self._c(lib().gf_dealias(self.h, int(pq), _ptr(v1), _ptr(v2), _ptr(out)))
        return out

    def dss(self, u, op="sum"):
        ops = {"sum": 0, "smooth": 1, "max": 2, "min": 3}
        self._c(lib().gf_dss(self.h, ops[op], _ptr(u)))
        return u

    def imult(self):
        out = self.new()
        self._c(lib().gf_imult(self.h, _ptr(out)))
        return out

    def advect(self, p, us, out=None):
        out = self.new() if out is None else out
        self._c(lib().gf_advect(self.h, _ptr(p), len(us), _ptrs(us), _ptr(out)))
        return out

    def burgers_rhs(self, us, nu, mode=0):
        # one output per component for the nonlinear (0) and heat (2) modes,
        # a single output for pure advection (1)
        nout = 1 if mode == 1 else len(us)
        outs = [self.new() for _ in range(nout)]
        self._c(lib().gf_burgers_rhs(self.h, mode, nu, len(us), _ptrs(us), _ptrs(outs)))
        return outs

    def burgers_step(self, us, nu, dt, t=0.0, mode=0, order=4, check_finite=False):
        self._c(lib().gf_burgers_step(self.h, mode, nu, order, len(us), t, dt, _ptrs(us),
                                      int(check_finite)))
        return us

    def burgers_steps(self, us, nu, dt, nsteps, t=0.0, mode=0, order=4, precision=0):
        """nsteps RK steps in place; precision 1 selects the mixed FP32/FP64
        stepper (pure advection on embedded meshes, gf_burgers_steps_mixed)."""
        fn = lib().gf_burgers_steps_mixed if precision else lib().gf_burgers_steps
        self._c(fn(self.h, mode, nu, order, len(us), t, dt, nsteps, _ptrs(us)))
        return us

    def burgers_rhs_adjoint(self, us, ws, lams, nu):
        """lams[k] += (dF/du)^T w at the state us (nonlinear mode)."""
        self._c(lib().gf_burgers_rhs_adjoint(self.h, nu, len(us), _ptrs(us), _ptrs(ws),
                                             _ptrs(lams)))
        return lams

    def burgers_step_adjoint(self, un, lams, nu, dt, order=4):
        """lams (dJ/du_{n+1}) becomes dJ/du_n for one step from the state un."""
        self._c(lib().gf_burgers_step_adjoint(self.h, nu, order, len(un), dt, _ptrs(un),
                                              _ptrs(lams)))
        return lams

    def burgers_adjoint(self, u0, lams, nu, dt, nsteps, order=4, ckpt=16, uT=None,
                        gio_prefix=None):
        """lams (dJ/du(T)) becomes dJ/du(0) for nsteps steps from u0; uT, when
        given, receives u(T).  gio_prefix: keep the checkpoints as GIO state
        files <prefix>.<c>.gio instead of device memory."""
        if gio_prefix is None:
            self._c(lib().gf_burgers_adjoint(self.h, nu, order, len(u0), dt, nsteps, ckpt,
                                             _ptrs(u0), _ptrs(lams),
                                             _ptrs(uT) if uT is not None else None))
        else:
            self._c(lib().gf_burgers_adjoint_gio(self.h, nu, order, len(u0), dt, nsteps, ckpt,
                                                 _ptrs(u0), _ptrs(lams),
                                                 _ptrs(uT) if uT is not None else None,
                                                 str(gio_prefix).encode()))
        return lams

    def set_segments(self, elem_offsets):
        """gf_mesh_set_segments: the mesh is a union of independent members
        (element ranges elem_offsets[k] .. elem_offsets[k+1]); CG solves then
        give every member its own scalars and stopping test."""
        eo = np.ascontiguousarray(elem_offsets, dtype=np.int64)
        self._c(lib().gf_mesh_set_segments(self.h, int(eo.size - 1),
                                           eo.ctypes.data_as(_i64p) if eo.size > 1 else None))
        self.nseg = int(eo.size - 1)

    def member_results(self):
        """Per-member (iterations, final norm, initial norm) of the last CG
        solve on a batched mesh."""
        n = self.nseg
        it = np.zeros(n, dtype=np.int32)
        rn, r0 = np.zeros(n), np.zeros(n)
        self._c(lib().gf_cg_member_results(self.h, n, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           _hp(rn), _hp(r0)))
        return it, rn, r0

    def time_kernel(self, which, reps=20):
        """Average ms per launch of one kernel, timed with HIP events."""
        ms = ctypes.c_double(0)
        self._c(lib().gf_time_kernel(self.h, which.encode(), reps, ctypes.byref(ms)))
        return ms.value

    def reserve(self, what, nstate=0, cg_maxit=0, adj_nsteps=0, adj_ckpt=0, adj_order=4):
        """gf_mesh_reserve: allocate the workspace of the selected entry
        points now (what: GF_RESERVE_* flags, GF_RESERVE_FREEZE to forbid
        later allocation)."""
        prm = _capi.ReserveParams(nstate, cg_maxit, adj_nsteps, adj_ckpt, adj_order)
        self._c(lib().gf_mesh_reserve(self.h, int(what), ctypes.byref(prm)))

    def helmholtz_diag_inv(self, p=1.0, q=None):
        out = self.new()
        qf, qs = _q_args(q)
        self._c(lib().gf_helmholtz_diag_inv(self.h, float(p), qs, _ptr(qf), _ptr(out)))
        return out

    def cg_solve(self, b, x=None, xb=None, p=1.0, q=None, maxit=512, tol=1e-6, normtype=0,
                 precond=0, precision=0, inner_tol=0.0, raise_on_error=False):
        x = self.new() if x is None else x
        prm = CGParams()
        prm.maxit = maxit
        prm.tol = tol
        prm.normtype = normtype
        prm.precond = precond
        prm.precision = precision
        prm.inner_tol = inner_tol
        prm.p_scalar = float(p)
        prm.p_field = None
        qf, qs = _q_args(q)
        prm.q_field = _ptr(qf)
        prm.q_scalar = qs
        hist = np.zeros(max(maxit, 1) + 8)
        prm.resid = hist.ctypes.data_as(_dp)
        res = CGResult()
        rc = lib().gf_cg_solve(self.h, ctypes.byref(prm), _ptr(b), _ptr(xb), _ptr(x),
                               ctypes.byref(res))
        if rc not in (0, _capi.GF_ERR_NOCONVERGE) or (rc and raise_on_error):
            self._c(rc)
        nh = res.iters if precision == 0 else res.outer + 1
        return x, dict(status=res.status, iters=res.iters, outer=res.outer, rnorm=res.rnorm,
                       rnorm0=res.rnorm0, resid=hist[: max(nh, 0)])
