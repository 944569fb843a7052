"""3-D GE_DEFORMED elements: a terrain-following box of hexahedra with the
geometry of GElem_base::dogeom3d / Jac / inv (src/cdg/sem/gelem_base.cpp:
972-1048, 1148-1168, 1268-1283) and the metric Helmholtz of def_init/def_prod
(src/cdg/sem/ghelmholtz.ipp:108-191, 494-514).

The reference's 3-D def_init runs both loops over all (i, j) while G(i,j) and
G(j,i) share one vector, so every off-diagonal metric vector receives its
terms twice; the product keeps that.

CPU: the product's generator (gf_terrain_box_grid3d) is bit-equal to the
oracle's restatement, and the oracle is pinned by what the mathematics fixes:
  * flat terrain is the regular box (Jacobian L^3/8, dXi/dX = diag(2/L)) and
    its deformed Helmholtz equals the regular-element Helmholtz;
  * the volume is 1 - amp/4 for the unit box;
  * the 3x3 inverse is the true one: the discrete chain rule returns
    dx_k/dx_l = delta_kl at every node;
  * the bilinear form: u^T L u = sum_q W J |grad u|^2 + X, with X the doubled
    off-diagonal terms sum_q W J sum_{i != j} (D_i u)(D_j u) g_ij,
    g_ij = sum_k dXi_i/dx_k dXi_j/dx_k, computed independently from the
    oracle's derivatives and metrics; X is not zero on sloped terrain.
GPU: mass, Helmholtz, derivatives, advection, DSS and RK4 Burgers steps on
the deformed mesh against the oracle at 1e-12; FP64 CG against the oracle.
"""
import numpy as np
import pytest

import oracle
import geoflow_amd as ga
from conftest import relerr

CASE = (3, [2, 3, 2], 0.3)


def test_terrain3_grid_matches_oracle():
    p, ne, amp = CASE
    og = oracle.Grid.terrain3(p, ne, amp)
    ig = ga.TerrainBoxGrid3D(p, ne, amp=amp)
    for d, cn in enumerate("xyz"):
        assert np.array_equal(ig.xyz[d], og.get(cn))
    assert np.array_equal(ig.node_jac, og.get("jac"))
    for i in range(3):
        for j in range(3):
            assert np.array_equal(ig.node_dxidx[i][j], og.get("dxidx_%d%d" % (i, j)))
    offs, mem = og.ggfx()
    assert np.array_equal(offs, ig.ggfx_offs) and np.array_equal(mem, ig.ggfx_members)
    assert ig.minnodedist == og.minnodedist


def test_terrain3_flat_is_the_regular_box():
    p, ne = 3, [2, 3, 2]
    og = oracle.Grid.terrain3(p, ne, 0.0)
    ob = oracle.Grid.box(3, p, ne, bdy=[6, 6, 6, 6, 0, 0])
    L = [1.0 / q for q in ne]
    assert np.abs(og.get("jac") - L[0] * L[1] * L[2] / 8).max() <= 1e-13 * L[0] * L[1] * L[2]
    for i in range(3):
        for j in range(3):
            want = 2.0 / L[i] if i == j else 0.0
            assert np.abs(og.get("dxidx_%d%d" % (i, j)) - want).max() <= 1e-12 * 2.0 / L[0]
    u = np.random.default_rng(1).standard_normal(og.ndof)
    assert relerr(og.helmholtz(u, 0.7, np.full(og.ndof, 1.3)),
                  ob.helmholtz(u, 0.7, np.full(ob.ndof, 1.3))) <= 1e-12


def test_terrain3_volume_and_chain_rule():
    p, ne, amp = CASE
    og = oracle.Grid.terrain3(p, ne, amp)
    assert abs(og.volume - (1.0 - amp / 4)) <= 1e-12
    X = [og.get(c) for c in "xyz"]
    for k in range(3):
        for l in range(3):
            d = og.deriv(X[k], l + 1)
            assert np.abs(d - (1.0 if k == l else 0.0)).max() <= 1e-11, (k, l)


def test_terrain3_bilinear_form_has_doubled_off_diagonals():
    p, ne, amp = CASE
    og = oracle.Grid.terrain3(p, ne, amp)
    u = og.dss(np.random.default_rng(4).standard_normal(og.ndof), "smooth")
    Du = [og.grefderiv(u, i + 1) for i in range(3)]
    dx = [[og.get("dxidx_%d%d" % (i, k)) for k in range(3)] for i in range(3)]
    WJ = og.mass_apply(np.ones(og.ndof))  # W0 W1 W2 Jac per node
    grad = [sum(Du[i] * dx[i][k] for i in range(3)) for k in range(3)]
    e_true = np.sum(WJ * sum(gk * gk for gk in grad))
    X = 0.0
    for i in range(3):
        for j in range(3):
            if i != j:
                gij = sum(dx[i][k] * dx[j][k] for k in range(3))
                X += np.sum(WJ * Du[i] * Du[j] * gij)
    e_ref = float(np.dot(u, og.helmholtz(u, 1.0, None)))
    assert abs(e_ref - (e_true + X)) <= 1e-11 * e_true
    assert abs(X) > 1e-4 * e_true


# ---------------------------------------------------------------------------
# GPU: operators on the deformed mesh against the oracle
# ---------------------------------------------------------------------------
TOL = 1e-12


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.gpu
def test_terrain3_operators_and_steps_match_oracle(ctx):
    from geoflow_amd.device import Mesh
    p, ne, amp = CASE
    print("terrain3 operators", flush=True)
    og = oracle.Grid.terrain3(p, ne, amp)
    ig = ga.TerrainBoxGrid3D(p, ne, amp=amp)
    m = Mesh(ctx, ig)
    u = np.random.default_rng(5).standard_normal(og.ndof)
    assert relerr(_host(m.mass_apply(_dev(u))), og.mass_apply(u)) <= TOL
    q = 1.0 + og.get("x")
    assert relerr(_host(m.helmholtz(_dev(u), 0.7, _dev(q))), og.helmholtz(u, 0.7, q)) <= TOL
    assert relerr(_host(m.helmholtz(_dev(u), 0.7, 1.3)),
                  og.helmholtz(u, 0.7, np.full(og.ndof, 1.3))) <= TOL
    for idir in (1, 2, 3):
        assert relerr(_host(m.deriv(_dev(u), idir)), og.deriv(u, idir)) <= TOL
    x, y, z = og.get("x"), og.get("y"), og.get("z")
    vel = [np.sin(2 * np.pi * x), np.cos(2 * np.pi * y), 0.1 * z * (1 - z)]
    assert relerr(_host(m.advect(_dev(u), [_dev(v) for v in vel])), og.advect(u, vel)) <= TOL
    for op in ("sum", "smooth", "max", "min"):
        assert np.array_equal(_host(m.dss(_dev(u), op)), og.dss(u, op))
    st = [0.3 * v for v in vel]
    dst = [_dev(a) for a in st]
    for s in range(2):
        m.burgers_step(dst, 0.02, 1e-3)
        og.burgers_step(st, 0.02, 1e-3)
        for k in range(3):
            assert relerr(_host(dst[k]), st[k]) <= TOL, (s, k)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", [0, 1])
def test_cg_on_3d_deformed_mesh(ctx, precond):
    """GCG (FP64, without and with the Jacobi preconditioner, whose diagonal
    is first checked against the oracle's) on the 3-D deformed box with
    Dirichlet top and bottom: 25 fixed iterations against the oracle, then a
    FP64 and mixed-precision solves to rel-res 1e-10 whose assembled residuals
    are checked directly."""
    from geoflow_amd.device import Mesh
    p, ne, amp = CASE
    print("cg on terrain3", flush=True)
    og = oracle.Grid.terrain3(p, ne, amp)
    ig = ga.TerrainBoxGrid3D(p, ne, amp=amp)
    m = Mesh(ctx, ig)
    f = og.dss(np.random.default_rng(22).standard_normal(og.ndof), "smooth")
    b = og.mass_apply(f)
    xb = np.zeros(og.ndof)
    if precond:
        assert relerr(_host(m.helmholtz_diag_inv(0.1, 1.0)),
                      og.helmholtz_diag_inv(0.1, np.ones(og.ndof))) <= 1e-13
    xg, info = m.cg_solve(_dev(b), xb=_dev(xb), p=0.1, q=1.0, maxit=25, tol=1e-30, normtype=0,
                          precond=precond)
    xo, ret, it, res = og.cg_solve(b, xb=xb, p=0.1, q=np.ones(og.ndof), maxit=25, tol=1e-30,
                                   normtype=0, precond=precond)
    assert info["iters"] == it == 25
    assert relerr(_host(xg), xo) <= 1e-11
    assert relerr(info["resid"][:25], res[:25]) <= 1e-10
    xg, info = m.cg_solve(_dev(b), xb=_dev(xb), p=0.1, q=1.0, maxit=3000, tol=1e-10, normtype=4,
                          precond=precond)
    assert info["status"] == 0
    mask = og.get("mask")
    r = og.dss(b - og.helmholtz(_host(xg), 0.1, np.ones(og.ndof)), "sum") * mask
    bA = og.dss(b.copy(), "sum") * mask
    im = og.imult()
    assert np.sqrt(np.sum(r * r * im)) <= 1.01e-10 * np.sqrt(np.sum(bA * bA * im))
    x64 = _host(xg)
    xm, info = m.cg_solve(_dev(b), xb=_dev(xb), p=0.1, q=1.0, maxit=3000, tol=1e-10, normtype=4,
                          precond=precond, precision=1)
    assert info["status"] == 0
    r = og.dss(b - og.helmholtz(_host(xm), 0.1, np.ones(og.ndof)), "sum") * mask
    assert np.sqrt(np.sum(r * r * im)) <= 1.01e-10 * np.sqrt(np.sum(bA * bA * im))
    assert relerr(_host(xm), x64) <= 1e-8
