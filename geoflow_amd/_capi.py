"""ctypes declarations for include/geoflow_hip.h."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GF_OK = 0
GF_ERR_NOCONVERGE = 1
GF_ERR_ARG = 10
GF_ERR_UNSUPPORTED = 14
GF_BDY_DIRICHLET = 0
GF_BDY_PERIODIC = 6
GF_BDY_NONE = 7
GF_GE_REGULAR = 0
GF_GE_DEFORMED = 1
GF_GE_2DEMBEDDED = 2
GF_DSS_SUM, GF_DSS_SMOOTH, GF_DSS_MAX, GF_DSS_MIN = 0, 1, 2, 3
GF_NORM_INF, GF_NORM_EUC, GF_NORM_L2, GF_NORM_L1, GF_NORM_ASSEMBLED_L2 = 0, 1, 2, 3, 4

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class GeoflowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("geoflow_hip error %d: %s" % (code, msg))
        self.code = code


GF_MESH_BOX_GATHER = 1
GF_MESH_BOX_MASK = 2
GF_MESH_CONST_D = 4
GF_MESH_BOX_HALO = 8
GF_MESH_IMASS_ELEM = 16
GF_ERR_PRECOND = 3
GF_ERR_ALLOC = 12
GF_RESERVE_RK, GF_RESERVE_CG, GF_RESERVE_CG_MIXED, GF_RESERVE_ADJOINT = 1, 2, 4, 8
GF_RESERVE_FREEZE = 256


class ReserveParams(ctypes.Structure):
    _fields_ = [("nstate", ctypes.c_int), ("cg_maxit", ctypes.c_int),
                ("adj_nsteps", ctypes.c_int), ("adj_ckpt", ctypes.c_int),
                ("adj_order", ctypes.c_int)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("gtype", ctypes.c_int), ("N", ctypes.c_int * 3),
                ("nelems", ctypes.c_size_t), ("D", _dp * 3), ("W", _dp * 3),
                ("elem_dxidx", _dp), ("elem_jac", _dp), ("node_dxidx", _dp * 9),
                ("node_jac", _dp), ("xyz", _dp * 3), ("mask", _dp),
                ("ggfx_offs", _i64p), ("ggfx_members", _i64p),
                ("box_ne", ctypes.c_int * 3), ("box_periodic", ctypes.c_int * 3),
                ("box_device_maps", ctypes.c_int), ("box_count", ctypes.c_int),
                ("box_list", ctypes.POINTER(ctypes.c_int))]


class GioInfo(ctypes.Structure):
    _fields_ = [("ivers", ctypes.c_int), ("dim", ctypes.c_int), ("nelems", ctypes.c_uint64),
                ("gtype", ctypes.c_int), ("cycle", ctypes.c_uint64), ("time", ctypes.c_double),
                ("multivar", ctypes.c_int), ("nporder", ctypes.c_uint64)]


class CGParams(ctypes.Structure):
    _fields_ = [("maxit", ctypes.c_int), ("tol", ctypes.c_double), ("normtype", ctypes.c_int),
                ("precond", ctypes.c_int), ("precision", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("p_scalar", ctypes.c_double),
                ("p_field", _vp), ("q_field", _vp), ("resid", _dp), ("q_scalar", ctypes.c_double)]


class CGResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int), ("outer", ctypes.c_int),
                ("rnorm", ctypes.c_double), ("rnorm0", ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "gf_version": (ctypes.c_char_p, []),
    "gf_ctx_create": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "gf_ctx_destroy": (ctypes.c_int, [_vp]),
    "gf_ctx_set_stream": (ctypes.c_int, [_vp, _vp]),
    "gf_ctx_synchronize": (ctypes.c_int, [_vp]),
    "gf_ctx_last_error": (ctypes.c_char_p, [_vp]),
    "gf_gll_basis": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
    "gf_box_grid": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_dp),
                                   ctypes.POINTER(_dp), _dp, _dp, _dp, _dp]),
    "gf_icos_grid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_dp),
                                    ctypes.POINTER(_dp), _dp, _dp]),
    "gf_deformed_geometry2d": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_dp),
                                              ctypes.c_int, ctypes.POINTER(_dp), _dp]),
    "gf_gio_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(GioInfo),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(_dp), ctypes.c_int, ctypes.c_size_t]),
    "gf_gio_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(GioInfo),
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_size_t,
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                   ctypes.POINTER(_dp), ctypes.c_int, ctypes.c_size_t]),
    "gf_deformed_geometry3d": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(_dp),
                                              ctypes.POINTER(_dp), _dp]),
    "gf_terrain_box_grid3d": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                             ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp,
                                             ctypes.POINTER(_dp), _dp, _dp]),
    "gf_terrain_box_grid": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, _dp,
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_double,
                                           ctypes.c_int, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                           _dp, ctypes.POINTER(_dp), _dp, _dp]),
    "gf_ggfx_build": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_dp),
                                     ctypes.c_double, ctypes.c_size_t, _i64p, _i64p,
                                     ctypes.POINTER(ctypes.c_size_t)]),
    "gf_mesh_create": (ctypes.c_int, [_vp, ctypes.POINTER(MeshDesc), ctypes.POINTER(_vp)]),
    "gf_mesh_destroy": (ctypes.c_int, [_vp]),
    "gf_mesh_ndof": (ctypes.c_size_t, [_vp]),
    "gf_mesh_flags": (ctypes.c_int, [_vp]),
    "gf_mesh_get": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "gf_mesh_reserve": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ReserveParams)]),
    "gf_malloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "gf_free": (ctypes.c_int, [_vp, _vp]),
    "gf_memcpy_h2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "gf_memcpy_d2h": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "gf_grefderiv": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "gf_grefderivs": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "gf_grefdiv": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), _vp]),
    "gf_deriv": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "gf_grad": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "gf_mass_apply": (ctypes.c_int, [_vp, _vp, _vp]),
    "gf_imass_apply": (ctypes.c_int, [_vp, _vp, _vp]),
    "gf_integrate": (ctypes.c_int, [_vp, _vp, _dp]),
    "gf_dealias": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "gf_mesh_set_segments": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "gf_cg_member_results": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            _dp, _dp]),
    "gf_box_ggfx_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), _vp, _vp,
                                          ctypes.POINTER(ctypes.c_size_t)]),
    "gf_integrate_noalloc": (ctypes.c_int, [_vp, _vp, _vp]),
    "gf_peak_probe": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "gf_helmholtz_apply": (ctypes.c_int, [_vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp,
                                          _vp]),
    "gf_helmholtz_apply_f32": (ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]),
    "gf_dss": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "gf_imult": (ctypes.c_int, [_vp, _vp]),
    "gf_advect": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(_vp), _vp]),
    "gf_burgers_rhs": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                      ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "gf_burgers_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(_vp), ctypes.c_int]),
    "gf_burgers_steps": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(_vp)]),
    "gf_burgers_steps_mixed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(_vp)]),
    "gf_burgers_rhs_adjoint": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int,
                                              ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp)]),
    "gf_burgers_step_adjoint": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, ctypes.POINTER(_vp),
                                               ctypes.POINTER(_vp)]),
    "gf_burgers_adjoint": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp)]),
    "gf_burgers_adjoint_gio": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                              ctypes.POINTER(_vp), ctypes.c_char_p]),
    "gf_cg_solve": (ctypes.c_int, [_vp, ctypes.POINTER(CGParams), _vp, _vp, _vp,
                                   ctypes.POINTER(CGResult)]),
    "gf_helmholtz_diag_inv": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "gf_time_kernel": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]),
}


def lib_path():
    return os.path.join(HERE, "libgeoflow_hip.so")


def lib():
    """Load libgeoflow_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libgeoflow_hip.so is not built: run __graft_entry__.build() "
                               "or python geoflow_amd/build.py")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, ctx=None):
    if rc != GF_OK:
        msg = ""
        if ctx is not None:
            m = lib().gf_ctx_last_error(ctx)
            msg = m.decode() if m else ""
        raise GeoflowError(rc, msg)
    return rc
