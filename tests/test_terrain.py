"""2-D GE_DEFORMED elements: a terrain-following box (the coordinates
GGrid::add_terrain installs, src/cdg/grid/ggrid.ipp:1486-1548, for an
analytic terrain) with the deformed-element geometry of GElem_base::dogeom2d,
Jac and inv (src/cdg/sem/gelem_base.cpp:877-955, 1121-1136, 1249-1266) and
the metric Helmholtz of def_init/def_prod (src/cdg/sem/ghelmholtz.ipp:108-191,
415-519).

CPU: the product's generator (gf_terrain_box_grid) is bit-equal to the
oracle's restatement (coordinates, Jacobian, dXi/dX up to last-bit noise in
near-zero entries, GGFX groups), and the
oracle is pinned by what the geometry fixes:
  * amp = 0 is the regular box: Jacobian L0 L1 / 4 and dXi/dX = diag(2/L);
  * the area is Lx H - amp Lx / 2 (the terrain h(x) = amp (1 - cos)/2 removes
    amp Lx / 2 under the bottom);
  * with the true inverse the discrete chain rule reproduces the
    coordinates: d x'/dx = d y'/dy = 1 and d x'/dy = d y'/dx = 0 at every node,
    and the assembled Laplacian of the linear field y' vanishes at every
    interior node (the 2-D metric identities of the collocated geometry);
  * the reference's 2x2 inverse is the transpose of the true one (its (0,1)
    entry is -R10/jac): with it d y'/dx equals the terrain slope term
    y'_xi / x_xi instead of 0 -- a reference quirk that the product keeps by
    default (reference_inverse=True) and that these pins document.
GPU: mass, Helmholtz, derivatives, advection, DSS, the Burgers RHS and RK4
steps on the deformed mesh against the oracle at 1e-12, for both inverses;
CG (FP64) on the deformed and on the sphere mesh against the oracle.
"""
import numpy as np
import pytest

import oracle
import geoflow_amd as ga
from conftest import relerr

CASES = [(4, [4, 3], 0.2), (6, [3, 2], 0.35)]


def _ids(c):
    return "p%d_%dx%d_a%g" % (c[0], c[1][0], c[1][1], c[2])


@pytest.mark.parametrize("ref", [True, False], ids=["refinv", "trueinv"])
@pytest.mark.parametrize("c", CASES, ids=_ids)
def test_terrain_grid_matches_oracle(c, ref):
    p, ne, amp = c
    og = oracle.Grid.terrain(p, ne, amp, ref_inverse=ref)
    ig = ga.TerrainBoxGrid(p, ne, amp=amp, reference_inverse=ref)
    assert ig.ndof == og.ndof
    for d, cn in enumerate("xy"):
        assert np.array_equal(ig.xyz[d], og.get(cn))
    assert np.array_equal(ig.node_jac, og.get("jac"))
    # dXi/dX bit for bit, except that the off-diagonal entries of a flat
    # direction are cancellation noise (~1e-14 of the diagonal) whose last
    # bit can differ between the two sums
    scale = max(np.abs(ig.node_dxidx[0][0]).max(), np.abs(ig.node_dxidx[1][1]).max())
    for i in range(2):
        for j in range(2):
            a, b = ig.node_dxidx[i][j], og.get("dxidx_%d%d" % (i, j))
            if i == j:
                assert np.array_equal(a, b)
            else:
                assert np.abs(a - b).max() <= 1e-15 * scale
    assert ig.minnodedist == og.minnodedist
    offs, mem = og.ggfx()
    assert np.array_equal(offs, ig.ggfx_offs) and np.array_equal(mem, ig.ggfx_members)


def test_terrain_flat_is_the_regular_box():
    ne, P1 = [4, 3], (1.0, 1.5)
    og = oracle.Grid.terrain(5, ne, 0.0, P1=P1)
    L = [P1[0] / ne[0], P1[1] / ne[1]]
    assert np.abs(og.get("jac") - L[0] * L[1] / 4).max() <= 1e-13 * L[0] * L[1]
    for i in range(2):
        for j in range(2):
            v = og.get("dxidx_%d%d" % (i, j))
            want = 2.0 / L[i] if i == j else 0.0
            assert np.abs(v - want).max() <= 1e-12 * 2.0 / L[0]


@pytest.mark.parametrize("c", CASES, ids=_ids)
def test_terrain_area(c):
    p, ne, amp = c
    og = oracle.Grid.terrain(p, ne, amp, P1=(2.0, 1.0))
    area = og.integrate(np.ones(og.ndof))
    assert abs(area - (2.0 * 1.0 - amp * 2.0 / 2)) <= 1e-12


def _interior(og):
    """Unique interior nodes: not on a Dirichlet face (mask 1)."""
    return og.get("mask") == 1.0


@pytest.mark.parametrize("c", CASES, ids=_ids)
def test_terrain_true_inverse_chain_rule_and_laplacian(c):
    p, ne, amp = c
    og = oracle.Grid.terrain(p, ne, amp, ref_inverse=False)
    x, y = og.get("x"), og.get("y")
    # x' jumps across the periodic seam; the chain rule is element-local
    assert np.abs(og.deriv(x, 1) - 1.0).max() <= 1e-11
    assert np.abs(og.deriv(x, 2)).max() <= 1e-11
    assert np.abs(og.deriv(y, 1)).max() <= 1e-11
    assert np.abs(og.deriv(y, 2) - 1.0).max() <= 1e-11
    Ly = og.dss(og.helmholtz(y, 1.0, None), "sum")
    inner = _interior(og)
    assert np.abs(Ly[inner]).max() <= 1e-11 * np.abs(Ly).max()


@pytest.mark.parametrize("c", CASES, ids=_ids)
def test_terrain_reference_inverse_is_transposed(c):
    """The reference's dXi/dX: d y'/dx is the slope term y'_xi / x_xi, not 0,
    and the assembled Laplacian of y' no longer vanishes inside."""
    p, ne, amp = c
    og = oracle.Grid.terrain(p, ne, amp, ref_inverse=True)
    ot = oracle.Grid.terrain(p, ne, amp, ref_inverse=False)
    y = og.get("y")
    # the off-diagonal entries are exchanged, the diagonal is the same
    assert np.array_equal(og.get("dxidx_01"), ot.get("dxidx_10"))
    assert np.array_equal(og.get("dxidx_10"), ot.get("dxidx_01"))
    assert np.array_equal(og.get("dxidx_00"), ot.get("dxidx_00"))
    assert np.abs(og.get("dxidx_01")).max() > 0.1 * amp
    r10 = og.grefderiv(y, 1)  # dy'/dxi
    r00 = og.grefderiv(og.get("x"), 1)
    slope = r10 / r00
    dydx = og.deriv(y, 1)
    assert np.abs(dydx - slope).max() <= 1e-11 * max(1.0, np.abs(slope).max())
    assert np.abs(slope).max() > 0.1 * amp
    Ly = og.dss(og.helmholtz(y, 1.0, None), "sum")
    assert np.abs(Ly[_interior(og)]).max() > 1e-6 * np.abs(Ly).max()


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
@pytest.mark.parametrize("ref", [True, False], ids=["refinv", "trueinv"])
def test_terrain_operators_match_oracle(ctx, ref):
    from geoflow_amd.device import Mesh
    p, ne, amp = CASES[1]
    print("terrain operators ref=%s" % ref, flush=True)
    og = oracle.Grid.terrain(p, ne, amp, ref_inverse=ref)
    ig = ga.TerrainBoxGrid(p, ne, amp=amp, reference_inverse=ref)
    m = Mesh(ctx, ig)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(og.ndof)
    assert relerr(_host(m.mass_apply(_dev(u))), og.mass_apply(u)) <= TOL
    assert relerr(_host(m.helmholtz(_dev(u), 0.7, 1.3)),
                  og.helmholtz(u, 0.7, np.full(og.ndof, 1.3))) <= TOL
    for idir in (1, 2):
        assert relerr(_host(m.deriv(_dev(u), idir)), og.deriv(u, idir)) <= TOL
    vel = [np.sin(2 * np.pi * og.get("x")), 0.3 + 0.1 * og.get("y")]
    assert relerr(_host(m.advect(_dev(u), [_dev(v) for v in vel])), og.advect(u, vel)) <= TOL
    for op in ("sum", "smooth", "max", "min"):
        assert np.array_equal(_host(m.dss(_dev(u), op)), og.dss(u, op))


@pytest.mark.gpu
def test_terrain_burgers_rk4_matches_oracle(ctx):
    from geoflow_amd.device import Mesh
    p, ne, amp = CASES[0]
    print("terrain burgers", flush=True)
    og = oracle.Grid.terrain(p, ne, amp)
    ig = ga.TerrainBoxGrid(p, ne, amp=amp)
    m = Mesh(ctx, ig)
    x, y = og.get("x"), og.get("y")
    st = [0.3 * np.sin(2 * np.pi * x) * (1.0 - y), 0.2 * np.cos(2 * np.pi * x) * y * (1.0 - y)]
    dst = [_dev(a) for a in st]
    ref = og.burgers_rhs(st, 0.02)
    from geoflow_amd.device import _ptrs
    from geoflow_amd._capi import lib
    outs = [m.new() for _ in range(2)]
    m._c(lib().gf_burgers_rhs(m.h, 0, 0.02, 2, _ptrs(dst), _ptrs(outs)))
    for k in range(2):
        assert relerr(_host(outs[k]), ref[k]) <= TOL, k
    for s in range(3):
        m.burgers_step(dst, 0.02, 1e-3)
        og.burgers_step(st, 0.02, 1e-3)
        for k in range(2):
            assert relerr(_host(dst[k]), st[k]) <= TOL, (s, k)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("mesh", ["terrain", "sphere"])
def test_cg_on_node_metric_meshes(ctx, mesh, precond):
    """GCG on 2-D meshes with per-node metrics (deformed and embedded): the
    Helmholtz (M + 0.1 L) on the terrain box with Dirichlet top and bottom
    (xb = 0) and on the icosahedral sphere (no boundary), without and with the
    Jacobi preconditioner (whose diagonal is first checked against the
    oracle's): 25 fixed iterations against the oracle, then FP64 and
    mixed-precision solves to rel-res 1e-10 whose assembled residuals are
    checked directly."""
    from geoflow_amd.device import Mesh
    import torch
    print("cg on %s" % mesh, flush=True)
    if mesh == "terrain":
        og = oracle.Grid.terrain(5, [4, 3], 0.25)
        ig = ga.TerrainBoxGrid(5, [4, 3], amp=0.25)
    else:
        og = oracle.Grid.icos(4, 1, radius=1.0)
        ig = ga.IcosGrid(4, 1, radius=1.0)
    m = Mesh(ctx, ig)
    f = np.random.default_rng(21).standard_normal(og.ndof)
    f = og.dss(f, "smooth")
    b = og.mass_apply(f)
    xb = np.zeros(og.ndof) if mesh == "terrain" else None
    if precond:
        assert relerr(_host(m.helmholtz_diag_inv(0.1, 1.0)),
                      og.helmholtz_diag_inv(0.1, np.ones(og.ndof))) <= 1e-13
    xg, info = m.cg_solve(_dev(b), xb=_dev(xb) if xb is not None else None, p=0.1, q=1.0,
                          maxit=25, tol=1e-30, normtype=0, precond=precond)
    xo, ret, it, res = og.cg_solve(b, xb=xb, p=0.1, q=np.ones(og.ndof), maxit=25, tol=1e-30,
                                   normtype=0, precond=precond)
    assert info["iters"] == it == 25
    assert relerr(_host(xg), xo) <= 1e-11
    assert relerr(info["resid"][:25], res[:25]) <= 1e-10
    xg, info = m.cg_solve(_dev(b), xb=_dev(xb) if xb is not None else None, p=0.1, q=1.0,
                          maxit=3000, tol=1e-10, normtype=4, precond=precond)
    assert info["status"] == 0
    r = og.dss(b - og.helmholtz(_host(xg), 0.1, np.ones(og.ndof)), "sum")
    if xb is not None:
        r *= og.get("mask")
    bA = og.dss(b.copy(), "sum") * (og.get("mask") if xb is not None else 1.0)
    im = og.imult()
    assert np.sqrt(np.sum(r * r * im)) <= 1.01e-10 * np.sqrt(np.sum(bA * bA * im))
    # FP32 recurrences with FP64 reliable updates reach the same FP64 criterion
    x64 = _host(xg)
    xm, info = m.cg_solve(_dev(b), xb=_dev(xb) if xb is not None else None, p=0.1, q=1.0,
                          maxit=3000, tol=1e-10, normtype=4, precond=precond, precision=1)
    assert info["status"] == 0
    r = og.dss(b - og.helmholtz(_host(xm), 0.1, np.ones(og.ndof)), "sum")
    if xb is not None:
        r *= og.get("mask")
    assert np.sqrt(np.sum(r * r * im)) <= 1.01e-10 * np.sqrt(np.sum(bA * bA * im))
    assert relerr(_host(xm), x64) <= 1e-8


# ---------------------------------------------------------------------------
# GGrid::add_terrain: Laplace solves for the coordinates
# ---------------------------------------------------------------------------
def _bump(x):
    return 0.2 * np.exp(-((x - 0.5) / 0.15) ** 2)


def _oracle_add_terrain(p, ne, terrain, tol=1e-13):
    """The oracle's add_terrain: gterrainSpecBox boundary data on the flat
    deformed grid, one oracle CG per coordinate, coordinates installed."""
    og = oracle.Grid.terrain(p, ne, 0.0, bdy=[0] * 4)
    x, y, mask = og.get("x"), og.get("y"), og.get("mask")
    bnd = mask == 0.0
    eps = 0.125 * og.minnodedist
    bottom = bnd & (np.abs(y - 0.0) <= eps)
    xb = [np.where(bnd, x, 0.0), np.where(bnd, y, 0.0) + np.where(bottom, terrain(x), 0.0)]
    new = []
    for j in range(2):
        xbj = og.dss(xb[j].copy(), "smooth")
        X, ret, it, res = og.cg_solve(np.zeros(og.ndof), xb=xbj, p=1.0, q=None, maxit=20000,
                                      tol=tol, normtype=0)
        assert ret == 0
        new.append(X)
    return og, xb, new


def test_add_terrain_oracle_pins():
    """Zero terrain reproduces the base coordinates (linear data are
    harmonic); with a terrain the solution takes the boundary data exactly
    and its assembled Laplacian vanishes at interior nodes."""
    og, xb, new = _oracle_add_terrain(4, [3, 3], lambda x: 0.0 * x)
    assert np.abs(new[0] - og.get("x")).max() <= 1e-11
    assert np.abs(new[1] - og.get("y")).max() <= 1e-11
    og, xb, new = _oracle_add_terrain(4, [3, 3], _bump)
    bnd = og.get("mask") == 0.0
    for j in range(2):
        assert np.array_equal(new[j][bnd], og.dss(xb[j].copy(), "smooth")[bnd])
        Lx = og.dss(og.helmholtz(new[j], 1.0, None), "sum")
        assert np.abs(Lx[~bnd]).max() <= 1e-10
    assert np.abs(new[1] - og.get("y")).max() > 0.05  # the terrain moved the nodes


@pytest.mark.gpu
def test_add_terrain_matches_oracle(ctx):
    from geoflow_amd.device import Mesh
    from geoflow_amd.terrain import add_terrain
    print("add_terrain", flush=True)
    p, ne = 5, [4, 3]
    grid, infos = add_terrain(ctx, p, ne, _bump, tol=1e-13)
    og, xb, new = _oracle_add_terrain(p, ne, _bump)
    for j in range(2):
        assert relerr(grid.xyz[j], new[j]) <= 1e-11, j
    og.set_coords(grid.xyz[0], grid.xyz[1])
    assert np.array_equal(grid.node_jac, og.get("jac"))
    for i in range(2):
        assert np.array_equal(grid.node_dxidx[i][i], og.get("dxidx_%d%d" % (i, i)))
    offs, mem = og.ggfx()
    assert np.array_equal(offs, grid.ggfx_offs) and np.array_equal(mem, grid.ggfx_members)
    # operators on the new grid
    m = Mesh(ctx, grid)
    u = np.random.default_rng(9).standard_normal(og.ndof)
    assert relerr(_host(m.helmholtz(_dev(u), 0.5, 1.0)),
                  og.helmholtz(u, 0.5, np.ones(og.ndof))) <= TOL
    assert abs(float(m.integrate(_dev(np.ones(og.ndof)))) - og.volume) <= 1e-13
