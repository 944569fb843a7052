"""Icosahedral sphere grid (GGridIcos 2-D, GE_2DEMBEDDED) and the operators
on it.

CPU: the product's generator (gf_icos_grid) against the oracle's restatement
(coordinates, Jacobian and dXi/dX bit for bit; GGFX groups bit for bit), and
the oracle pinned by what the geometry fixes:
  * cdg_mass.cpp:132-229: the integral of 1 over the level-2 bisection sphere
    at p=4 is 4 pi R^2 to relative error < 1e-10 (the test's criterion);
  * every node lies on the sphere; element counts 60 4^L (bisection) and
    60 (L+1)^2 (Lagrangian refinement);
  * Euler's formula for the quad mesh fixes the number of unique nodes, which
    the GGFX multiplicities must sum to: 3T(N-2)^2 + 6T(N-2) + 3T + 2 for T
    triangles;
  * z is a degree-1 spherical harmonic on the unit sphere: the assembled
    Laplacian gives L z = 2 M z up to discretisation error, and L 1 = 0.
GPU: mass, Helmholtz, derivatives, advection and DSS on the sphere, RK4
steps of pure advection, advection with viscosity at RK orders 1-4 (the
fused embedded stage kernel) and nonlinear Burgers (with the tangent-plane
projection), against the oracle at 1e-12.
"""
import numpy as np
import pytest

import oracle
import geoflow_amd as ga
from conftest import relerr

LEVELS = [(4, 1, 0), (5, 1, 0), (3, 2, 0), (4, 2, 1)]


def _ids(c):
    return "p%d_L%d_%s" % (c[0], c[1], "bis" if c[2] == 0 else "lag")


@pytest.mark.parametrize("c", LEVELS, ids=_ids)
def test_icos_grid_matches_oracle(c):
    p, lev, ref = c
    og = oracle.Grid.icos(p, lev, radius=1.5, refine=ref)
    ig = ga.IcosGrid(p, lev, radius=1.5, refine=ref)
    assert ig.ndof == og.ndof and ig.nelems == og.nelems
    for d, cn in enumerate("xyz"):
        assert np.array_equal(ig.xyz[d], og.get(cn))
    assert np.array_equal(ig.node_jac, og.get("jac"))
    for i in range(2):
        for j in range(3):
            assert np.array_equal(ig.node_dxidx[i][j], og.get("dxidx_%d%d" % (i, j)))
    assert ig.minnodedist == og.minnodedist
    offs, mem = og.ggfx()
    assert np.array_equal(offs, ig.ggfx_offs)
    assert np.array_equal(mem, ig.ggfx_members)


def test_icos_sphere_area_cdg_mass():
    """cdg_mass.cpp: level 2, GICOS_BISECTION, p=4; |A - 4 pi R^2| / 4 pi R^2 < 1e-10."""
    for R in (1.0, 6.37122e6):
        og = oracle.Grid.icos(4, 2, radius=R)
        area = og.integrate(np.ones(og.ndof))
        assert abs(area - 4 * np.pi * R * R) / (4 * np.pi * R * R) < 1e-10


@pytest.mark.parametrize("c", LEVELS, ids=_ids)
def test_icos_counts_radius_and_multiplicity(c):
    p, lev, ref = c
    og = oracle.Grid.icos(p, lev, radius=2.0, refine=ref)
    T = 20 * 4 ** lev if ref == 0 else 20 * (lev + 1) ** 2
    assert og.nelems == 3 * T
    r = np.sqrt(og.get("x") ** 2 + og.get("y") ** 2 + og.get("z") ** 2)
    assert np.abs(r - 2.0).max() <= 4 * np.finfo(float).eps * 2.0
    N = p + 1
    unique = 3 * T * (N - 2) ** 2 + 6 * T * (N - 2) + 3 * T + 2
    assert abs(og.imult().sum() - unique) < 1e-6 * unique
    offs, mem = og.ggfx()
    assert np.diff(offs).max() <= 6  # GGridIcos::max_duplicates (2-D)


def test_icos_laplacian_spherical_harmonic():
    og = oracle.Grid.icos(6, 2, radius=1.0)
    z = og.get("z")
    Lz = og.dss(og.helmholtz(z, 1.0, None), "sum")
    Mz = og.dss(og.mass_apply(z), "sum")
    assert np.abs(Lz - 2.0 * Mz).max() / np.abs(Mz).max() < 1e-6
    L1 = og.dss(og.helmholtz(np.ones(og.ndof), 1.0, None), "sum")
    # round-off only: assembled Lz entries are ~1e-3, L1 entries ~1e-15
    assert np.abs(L1).max() < 1e-9 * np.abs(Lz).max()


# ---------------------------------------------------------------------------
# GPU: operators on the sphere against the oracle
# ---------------------------------------------------------------------------
TOL = 1e-12


def _make(ctx, p=5, lev=1):
    from geoflow_amd.device import Mesh
    print("icos mesh p=%d level=%d" % (p, lev), flush=True)
    og = oracle.Grid.icos(p, lev, radius=1.0)
    ig = ga.IcosGrid(p, lev, radius=1.0)
    return og, ig, Mesh(ctx, ig)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _bell(x, y, z, lat0=0.3, lon0=1.1, R=1.0):
    lat, lon = np.arcsin(np.clip(z / R, -1, 1)), np.arctan2(y, x)
    r = R * np.arccos(np.clip(np.sin(lat0) * np.sin(lat) + np.cos(lat0) * np.cos(lat) *
                              np.cos(lon - lon0), -1, 1))
    return np.where(r < R / 3, 0.5 * (1 + np.cos(3 * np.pi * r / R)), 0.0)


def _rotation(x, y, z, alpha=np.pi / 4, w=2 * np.pi):
    ax = np.array([-np.sin(alpha), 0.0, np.cos(alpha)]) * w
    return [ax[1] * z - ax[2] * y, ax[2] * x - ax[0] * z, ax[0] * y - ax[1] * x]


@pytest.mark.gpu
def test_icos_operators_match_oracle(ctx):
    og, ig, m = _make(ctx)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(og.ndof)
    assert relerr(_host(m.mass_apply(_dev(u))), og.mass_apply(u)) <= TOL
    assert relerr(_host(m.helmholtz(_dev(u), 0.7, 1.3)), og.helmholtz(u, 0.7, np.full(og.ndof, 1.3))) <= TOL
    for idir in (1, 2, 3):
        assert relerr(_host(m.deriv(_dev(u), idir)), og.deriv(u, idir)) <= TOL
    vel = _rotation(*ig.xyz)
    assert relerr(_host(m.advect(_dev(u), [_dev(v) for v in vel])), og.advect(u, vel)) <= TOL
    for op in ("sum", "smooth", "max", "min"):
        assert np.array_equal(_host(m.dss(_dev(u), op)), og.dss(u, op))


@pytest.mark.gpu
def test_icos_pure_advection_rk4_matches_oracle(ctx):
    og, ig, m = _make(ctx)
    h = _bell(*ig.xyz)
    vel = _rotation(*ig.xyz)
    st = [h.copy()] + [v.copy() for v in vel]
    dst = [_dev(a) for a in st]
    for s in range(5):
        m.burgers_step(dst, 0.0, 1e-3, mode=1)
        og.burgers_step(st, 0.0, 1e-3, mode=1)
        assert relerr(_host(dst[0]), st[0]) <= TOL, s


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_icos_viscous_advection_orders_match_oracle(ctx, order):
    """The fused embedded advection stage with a viscous term, every RK order
    (Euler's single stage smooths its input; RK2-4 carry the stage input and
    the state accumulation inside the stage kernels), and a state whose
    coincident nodes disagree, which only the reference's smoothing of the
    last stage's input reconciles."""
    og, ig, m = _make(ctx, p=4, lev=2)
    h = _bell(*ig.xyz, lat0=-0.2, lon0=2.0)
    h = h + 1e-3 * np.random.default_rng(order).standard_normal(h.size)
    vel = _rotation(*ig.xyz, alpha=0.3)
    st = [h.copy()] + [v.copy() for v in vel]
    dst = [_dev(a) for a in st]
    for s in range(3):
        m.burgers_step(dst, 2e-3, 5e-4, mode=1, order=order)
        og.burgers_step(st, 2e-3, 5e-4, mode=1, order=order)
        assert relerr(_host(dst[0]), st[0]) <= TOL, s
        for k in range(1, 4):
            assert np.array_equal(_host(dst[k]), vel[k - 1])


@pytest.mark.gpu
def test_icos_nonlinear_burgers_rk4_matches_oracle(ctx):
    """Three velocity components on the sphere: step_b projects them onto the
    tangent plane (constrain2sphere, gmtk.ipp:2914-2980); dudt_impl forms the
    RHS of the first GDIM = 2 components only (gburgers.ipp:247)."""
    og, ig, m = _make(ctx, p=4)
    st = [0.2 * v for v in _rotation(*ig.xyz)]
    dst = [_dev(a) for a in st]
    # the RHS, with the third component's output poisoned beforehand
    outs = [m.new() for _ in range(3)]
    outs[2].fill_(7.0)
    from geoflow_amd.device import _ptrs
    from geoflow_amd._capi import lib
    m._c(lib().gf_burgers_rhs(m.h, 0, 0.01, 3, _ptrs(dst), _ptrs(outs)))
    ref = og.burgers_rhs(st, 0.01)
    for k in range(2):
        assert relerr(_host(outs[k]), ref[k]) <= TOL, k
    assert not _host(outs[2]).any() and not ref[2].any()
    for s in range(3):
        m.burgers_step(dst, 0.01, 1e-3)
        og.burgers_step(st, 0.01, 1e-3)
        for k in range(3):
            assert relerr(_host(dst[k]), st[k]) <= TOL, (s, k)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 4])
def test_icos_mixed_precision_advection_against_oracle(ctx, order):
    """gf_burgers_steps_mixed: the advected field and its stage fields in
    FP32 with the sphere's metric terms in FP64.  After 20 steps of a
    viscous cosine bell under tilted solid-body rotation (p=5, level 2) the
    FP32 state agrees with the FP64 oracle to 1e-5 of its maximum: FP32
    rounds each operation to 6e-8, and the steps carry those errors along
    and amplify them (forward Euler, which amplifies every advected mode,
    measures 5.6e-6 after 20 steps); the velocity is untouched; the FP64
    steps on the same mesh agree with the oracle to the FP64 tolerance."""
    og, ig, m = _make(ctx, p=5, lev=2)
    h = _bell(*ig.xyz, lat0=0.3, lon0=1.0)
    vel = _rotation(*ig.xyz, alpha=np.pi / 4)
    st = [h.copy()] + [v.copy() for v in vel]
    d64 = [_dev(a) for a in st]
    d32 = [_dev(a) for a in st]
    nu, dt, n = 5e-5, 2e-3, 20
    m.burgers_steps(d64, nu, dt, n, mode=1, order=order)
    m.burgers_steps(d32, nu, dt, n, mode=1, order=order, precision=1)
    for _ in range(n):
        og.burgers_step(st, nu, dt, mode=1, order=order)
    assert relerr(_host(d64[0]), st[0]) <= 1e-11
    err32 = np.abs(_host(d32[0]) - st[0]).max() / np.abs(st[0]).max()
    assert 0 < err32 <= 1e-5, err32
    for k in range(1, 4):
        assert np.array_equal(_host(d32[k]), vel[k - 1])


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3, 4])
def test_icos_multistep_equals_single_steps(ctx, order):
    """gf_burgers_steps on an embedded mesh runs its steps in one stepper
    call (the step's first state accumulation starts from the step-start
    copy): n steps in one call equal n one-step calls bit for bit, in FP64
    and in the mixed stepper, from a state whose coincident nodes disagree
    (which the first step must not smooth), and the FP64 steps agree with
    the oracle."""
    og, ig, m = _make(ctx, p=4, lev=2)
    h = _bell(*ig.xyz, lat0=0.4, lon0=-1.0)
    h = h + 1e-3 * np.random.default_rng(10 + order).standard_normal(h.size)
    vel = _rotation(*ig.xyz, alpha=0.7)
    st = [h.copy()] + [v.copy() for v in vel]
    nu, dt, n = 2e-3, 5e-4, 4
    for prec in (0, 1):
        one = [_dev(a) for a in st]
        many = [_dev(a) for a in st]
        m.burgers_steps(many, nu, dt, n, mode=1, order=order, precision=prec)
        for _ in range(n):
            m.burgers_steps(one, nu, dt, 1, mode=1, order=order, precision=prec)
        assert np.array_equal(_host(many[0]), _host(one[0])), prec
        if prec == 0:
            ref = [a.copy() for a in st]
            for _ in range(n):
                og.burgers_step(ref, nu, dt, mode=1, order=order)
            assert relerr(_host(many[0]), ref[0]) <= TOL


@pytest.mark.gpu
def test_mixed_steps_refuse_other_modes(ctx):
    from geoflow_amd import _capi
    og, ig, m = _make(ctx, p=4)
    st = [_dev(0.2 * v) for v in _rotation(*ig.xyz)]
    with pytest.raises(_capi.GeoflowError) as ei:
        m.burgers_steps(st, 0.01, 1e-3, 2, mode=0, precision=1)
    assert ei.value.code == _capi.GF_ERR_UNSUPPORTED
