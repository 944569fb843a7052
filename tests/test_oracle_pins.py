"""Pins for the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible mistake in the oracle -- a dropped
term, a wrong sign or index, a transposed operand, a wrong coefficient --
fails it: known-answer tests from the reference's own test suite, the
reference's compiled dmxm kernels, closed forms and brute force on tiny
inputs.  No check retypes the oracle's own formula.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

EPS = np.finfo(float).eps
HERE = os.path.dirname(os.path.abspath(__file__))
REF_DMXM = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_dmxm.so")


# ---------------------------------------------------------------------------
# GLL basis (gllbasis.ipp): closed forms and polynomial identities
# ---------------------------------------------------------------------------
def test_gll_closed_forms():
    xi, w, D = oracle.gll_basis(2)
    assert np.allclose(xi, [-1, 0, 1], atol=1e-15, rtol=0)
    assert np.allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15, rtol=0)
    xi, w, D = oracle.gll_basis(3)
    assert np.allclose(xi, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], atol=1e-15, rtol=0)
    assert np.allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], atol=1e-15, rtol=0)
    xi, w, D = oracle.gll_basis(4)
    r = np.sqrt(3 / 7)
    assert np.allclose(xi, [-1, -r, 0, r, 1], atol=1e-15, rtol=0)
    assert np.allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15, rtol=0)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6, 7, 8, 10, 12])
def test_gll_identities(p):
    xi, w, D = oracle.gll_basis(p)
    N = p + 1
    assert abs(w.sum() - 2.0) <= 10 * EPS
    assert np.allclose(xi, -xi[::-1], atol=2 * EPS, rtol=0)            # symmetric nodes
    assert np.all(np.diff(xi) > 0)
    assert abs(D[0, 0] + p * (p + 1) / 4) == 0.0                        # D(0,0) = -p(p+1)/4
    assert abs(D[p, p] - p * (p + 1) / 4) == 0.0
    for k in range(0, p + 1):                                           # D xi^k = k xi^(k-1)
        exact = k * xi ** (k - 1) if k > 0 else np.zeros(N)
        assert np.abs(D @ xi ** k - exact).max() <= 200 * EPS * max(1, p * p)
    # quadrature exact to degree 2p-1
    for k in range(0, 2 * p):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(w @ xi ** k - exact) <= 50 * EPS


# ---------------------------------------------------------------------------
# dmxm against the reference's own compiled kernels (dmxmp.F)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not os.path.exists(REF_DMXM), reason="oracle/_ref not built")
@pytest.mark.parametrize("nbi", [2, 5, 7, 8, 9, 13, 24])
def test_dmxm_matches_reference_fortran_bitwise(nbi):
    ref = ctypes.CDLL(REF_DMXM)
    f = getattr(ref, "dmxm%d_" % nbi)
    rng = np.random.default_rng(nbi)
    na, nc = 7, 11
    A = np.asfortranarray(rng.standard_normal((na, nbi)))
    B = np.asfortranarray(rng.standard_normal((nbi, nc)))
    C = np.zeros((na, nc), order="F")
    ci = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    f(A.ctypes.data_as(dp), ctypes.byref(ci(na)), B.ctypes.data_as(dp), ctypes.byref(ci(nc)),
      C.ctypes.data_as(dp), ctypes.byref(ci(nc)))
    ours = oracle.dmxm(A, B)
    assert np.array_equal(ours, C)


# ---------------------------------------------------------------------------
# GMTK tensor products: integer known-answer test of cdg_gmtk.cpp:77-268
# ---------------------------------------------------------------------------
def _kat_mats():
    np_ = 4
    N = [np_ + 1, np_ + 2, np_ + 3]
    Ds = []
    for n in N:
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        Ds.append((2.0 * i + N[0] * j).astype(np.float64))
    return N, Ds


def test_gmtk_kat_3d():
    N, Ds = _kat_mats()
    kk, jj, ii = np.meshgrid(np.arange(N[2]), np.arange(N[1]), np.arange(N[0]), indexing="ij")
    u = (2.0 * ii + N[0] * jj).reshape(-1)
    I = [np.eye(n) for n in N]
    ops = {1: np.kron(I[2], np.kron(I[1], Ds[0])), 2: np.kron(I[2], np.kron(Ds[1], I[0])),
           3: np.kron(Ds[2], np.kron(I[1], I[0]))}
    for idir in (1, 2, 3):
        y = oracle.grefderiv_raw(3, N, 1, Ds, idir, False, u)
        assert np.array_equal(y, ops[idir] @ u)
        yt = oracle.grefderiv_raw(3, N, 1, Ds, idir, True, u)
        assert np.array_equal(yt, ops[idir].T @ u)
    # D3 x D2 x D1 as the composition of the three (cdg_gmtk.cpp:235-268)
    y = u
    for idir in (1, 2, 3):
        y = oracle.grefderiv_raw(3, N, 1, Ds, idir, False, y)
    big = np.kron(Ds[2], np.kron(Ds[1], Ds[0]))
    assert np.array_equal(y, big @ u)


def test_gmtk_kat_2d():
    N, Ds = _kat_mats()
    N2 = N[:2]
    jj, ii = np.meshgrid(np.arange(N2[1]), np.arange(N2[0]), indexing="ij")
    u = (2.0 * ii + N[0] * jj).reshape(-1)
    ops = {1: np.kron(np.eye(N2[1]), Ds[0]), 2: np.kron(Ds[1], np.eye(N2[0]))}
    for idir in (1, 2):
        assert np.array_equal(oracle.grefderiv_raw(2, N2, 1, Ds[:2], idir, False, u), ops[idir] @ u)


# ---------------------------------------------------------------------------
# Derivatives: polynomial exactness (hackathon_driver.cpp:177-239; gtest_derivs)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("p", [3, 4, 7])
def test_deriv_polynomial_exactness(p):
    g = oracle.Grid.box(3, p, [3, 2, 2], P0=[0.0, -1.0, 0.5], P1=[1.5, 1.0, 2.0], bdy=[0] * 6)
    x, y, z = g.get("x"), g.get("y"), g.get("z")
    for (a, b, c) in [(p, 0, 0), (0, p, 1), (2, 2, 1), (1, p - 1, p)]:
        u = x ** a * y ** b * z ** c
        exact = [a * x ** max(a - 1, 0) * y ** b * z ** c,
                 b * x ** a * y ** max(b - 1, 0) * z ** c,
                 c * x ** a * y ** b * z ** max(c - 1, 0)]
        for d in range(3):
            du = g.deriv(u, d + 1)
            # inf-norm error relative to the derivative's scale (for an exactly
            # zero derivative, the scale is max|u| times the inverse element size)
            scale = max(np.abs(exact[d]).max(), np.abs(u).max() * 2.0 / 0.5)
            err = np.abs(du - exact[d]).max() / scale
            assert err <= 100 * EPS * max(1, p - 3), (a, b, c, d, err)


def test_hackathon_x2y2z_p4():
    """hackathon_driver.cpp:177-239 with hack_input.jsn: u = x^2 y^2 z on a
    fully periodic unit box of 100x100x1 elements, p=4, d/dx; error
    sqrt(integrate((da-du)^2)) / max|da| must be <= 100 eps (published
    2.01001e-14, docs/doxy/hackathon_20201201/compile_and_run.md:84-85;
    this oracle gives 1.29e-14)."""
    g = oracle.Grid.box(3, 4, [100, 100, 1])
    x, y, z = g.get("x"), g.get("y"), g.get("z")
    u = x ** 2 * y ** 2 * z
    da = 2 * x * y ** 2 * z
    du = g.deriv(u, 1)
    err = np.sqrt(g.integrate((da - du) ** 2)) / np.abs(da).max()
    assert err <= 100 * EPS
    assert 2e-15 < err < 1e-13  # same magnitude as the published figure


# ---------------------------------------------------------------------------
# GGFX: brute-force matching, closed-form multiplicities, doOp identities
# ---------------------------------------------------------------------------
def _brute_groups(g, tol):
    X = np.stack([g.get("xp"), g.get("yp"), g.get("zp") if g.dim == 3 else np.zeros(g.ndof)], 1)
    n = len(X)
    groups = [None] * n
    for i in range(n):
        if groups[i] is not None:
            continue
        lo, hi = X[i] - tol, X[i] + tol
        hit = np.all((lo <= X + tol) & (hi >= X - tol), axis=1)
        ids = np.flatnonzero(hit)
        for m in ids:
            groups[m] = ids
    return groups


@pytest.mark.parametrize("dim,p,ne,bdy", [(2, 3, [3, 2], [6] * 4), (2, 2, [2, 3], [0] * 4),
                                          (3, 2, [2, 2, 3], [6] * 6), (3, 3, [2, 1, 2], [0] * 6),
                                          (3, 2, [2, 3, 1], [6, 6, 0, 0, 6, 6])])
def test_ggfx_vs_bruteforce(dim, p, ne, bdy):
    g = oracle.Grid.box(dim, p, ne, bdy=bdy)
    offs, mem = g.ggfx()
    bf = _brute_groups(g, 0.001 * g.minnodedist)
    for i in range(g.ndof):
        assert np.array_equal(mem[offs[i]:offs[i + 1]], bf[i])


def test_multiplicity_closed_form_periodic():
    g = oracle.Grid.box(3, 4, [3, 3, 2])
    mult = np.diff(g.ggfx()[0])
    N = 5
    i = np.arange(g.ndof) % N
    j = (np.arange(g.ndof) // N) % N
    k = (np.arange(g.ndof) // (N * N)) % N
    onf = lambda a: ((a == 0) | (a == N - 1)).astype(int)
    expect = 2 ** (onf(i) + onf(j) + onf(k))
    assert np.array_equal(mult, expect)


def test_multiplicity_closed_form_nonperiodic():
    """gtest_mult.cpp:148-300 logic: boundary-reduced multiplicities."""
    ne = [3, 2]
    p = 3
    N = p + 1
    g = oracle.Grid.box(2, p, ne, bdy=[0] * 4)
    mult = np.diff(g.ggfx()[0])
    idx = np.arange(g.ndof)
    e = idx // (N * N)
    r = idx % (N * N)
    i, j = r % N, r // N
    ex, ey = e % ne[0], e // ne[0]

    def m1(a, ea, nea):
        if a == 0:
            return 2 if ea > 0 else 1
        if a == N - 1:
            return 2 if ea < nea - 1 else 1
        return 1
    expect = np.array([m1(i[q], ex[q], ne[0]) * m1(j[q], ey[q], ne[1]) for q in range(g.ndof)])
    assert np.array_equal(mult, expect)


def test_dss_identities():
    g = oracle.Grid.box(3, 3, [2, 3, 2])
    x, y, z = g.get("x"), g.get("y"), g.get("z")
    cont = np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y) + z * 0.0  # continuous, periodic
    assert np.allclose(g.dss(cont, "smooth"), cont, rtol=0, atol=4 * EPS)   # cdg_ggfx.cpp:217-225
    mult = np.diff(g.ggfx()[0])
    assert np.array_equal(g.dss(np.ones(g.ndof), "sum"), mult.astype(float))
    assert np.array_equal(g.imult(), 1.0 / mult)
    u = np.random.default_rng(0).standard_normal(g.ndof)
    mx, mn = g.dss(u, "max"), g.dss(u, "min")
    offs, mem = g.ggfx()
    for i in range(0, g.ndof, 7):
        grp = mem[offs[i]:offs[i + 1]]
        assert mx[i] == u[grp].max() and mn[i] == u[grp].min()
        s = 0.0
        for m in grp:          # std::accumulate from 0, ascending ids
            s = s + u[m]
        assert g.dss(u, "sum")[i] == s


# ---------------------------------------------------------------------------
# Mass and integrate
# ---------------------------------------------------------------------------
def test_mass_volume_and_inverse():
    g = oracle.Grid.box(3, 5, [2, 3, 2], P0=[0, 0, 0], P1=[2.0, 1.5, 0.5], bdy=[0] * 6)
    mass = g.get("mass")
    assert abs(mass.sum() - 1.5) <= 1e-13
    assert abs(g.volume - 1.5) <= 1e-13
    # imass * DSS_sum(mass) / mult = 1 (inverse of the averaged mass, gmass.ipp:75-82)
    mult = np.diff(g.ggfx()[0])
    assert np.allclose(g.get("imass") * g.dss(mass, "sum") / mult, 1.0, rtol=0, atol=8 * EPS)
    x = g.get("x")
    assert abs(g.integrate(x ** 3) - 0.75 * 2.0 ** 4 / 4 * 0.5 * 2) < 1e-13 * 10


# ---------------------------------------------------------------------------
# Helmholtz: SPD assembled operator (cdg_cg.cpp:213-245), L 1 = 0, energy
# ---------------------------------------------------------------------------
def _assembled(g, apply):
    offs, mem = g.ggfx()
    first = mem[offs[:-1]]
    uniq = np.unique(first)
    col = {u: k for k, u in enumerate(uniq)}
    n = len(uniq)
    A = np.zeros((n, n))
    for c, u in enumerate(uniq):
        e = np.zeros(g.ndof)
        e[mem[offs[u]:offs[u + 1]]] = 1.0
        y = g.dss(apply(e), "sum")
        A[:, c] = y[uniq]
    return A


def test_helmholtz_symmetric_and_nullspace():
    g = oracle.Grid.box(2, 4, [2, 2])
    A = _assembled(g, lambda e: g.helmholtz(e))
    assert np.abs(A - A.T).max() <= 100 * EPS * np.abs(A).max()
    assert np.abs(g.helmholtz(np.ones(g.ndof))).max() <= 1e-12
    ev = np.linalg.eigvalsh(A)
    assert ev[0] > -1e-12 and ev[1] > 1e-8  # one zero mode (periodic), rest positive
    g3 = oracle.Grid.box(3, 3, [2, 1, 2], bdy=[0] * 6)
    q = np.ones(g3.ndof)
    A3 = _assembled(g3, lambda e: g3.helmholtz(e, 0.1, q))
    assert np.abs(A3 - A3.T).max() <= 100 * EPS * np.abs(A3).max()
    assert np.linalg.eigvalsh(A3)[0] > 0


def test_helmholtz_energy_quadratic_form():
    """u^T L u = int |grad u|^2 for a polynomial u resolved exactly."""
    g = oracle.Grid.box(3, 4, [2, 2, 2], P1=[1.0, 2.0, 1.0], bdy=[0] * 6)
    x, y, z = g.get("x"), g.get("y"), g.get("z")
    u = x * y + z * z
    Lu = g.helmholtz(u)
    # int over [0,1]x[0,2]x[0,1] of y^2 + x^2 + 4 z^2 = 8/3 + 2/3 + 8/3
    assert abs(u @ Lu - (8 / 3 + 2 / 3 + 8 / 3)) <= 1e-12
    assert abs(u @ g.helmholtz(u, 0.25) - 0.25 * (8 / 3 + 2 / 3 + 8 / 3)) <= 1e-12


# ---------------------------------------------------------------------------
# CG: the cdg_cg.cpp Poisson test (2-D 4x4, p=6, Dirichlet, INF tol 1e-7)
# ---------------------------------------------------------------------------
def test_cg_poisson_reference_criteria():
    g = oracle.Grid.box(2, 6, [4, 4], bdy=[0] * 4)
    x, y = g.get("x"), g.get("y")
    f = 6.0 * x * y * (1.0 - y) - 2.0 * x ** 3
    ua = y * (1.0 - y) * x ** 3
    ub = np.where(np.abs(x - 1.0) <= 100 * EPS, y * (1.0 - y), 0.0)
    b = -g.mass_apply(f)
    u, ret, it, res = g.cg_solve(b, xb=ub, maxit=256, tol=1e-7, normtype=0)
    assert ret == 0
    err = g.integrate((u - ua) ** 2) / g.volume
    assert err < 1e-12


def test_cg_matches_dense_solve():
    g = oracle.Grid.box(3, 3, [2, 2, 1], bdy=[0] * 6)
    q = np.ones(g.ndof)
    f = np.random.default_rng(3).standard_normal(g.ndof)
    # GCG assembles b - A x itself (gcg.ipp:268-270): b is the local RHS
    b = g.mass_apply(f)
    x, ret, it, res = g.cg_solve(b, xb=np.zeros(g.ndof), p=0.1, q=q, maxit=500, tol=1e-14,
                                 normtype=4, precond=1)
    assert ret == 0
    offs, mem = g.ggfx()
    mask = g.get("mask")
    uniq = np.unique(mem[offs[:-1]])
    inner = uniq[mask[uniq] > 0]
    A = _assembled(g, lambda e: g.helmholtz(e, 0.1, q))
    pos = {u: k for k, u in enumerate(uniq)}
    sel = [pos[u] for u in inner]
    bA = g.dss(b, "sum")
    xs = np.linalg.solve(A[np.ix_(sel, sel)], bA[inner])
    assert np.abs(x[inner] - xs).max() <= 1e-11 * np.abs(xs).max()


# ---------------------------------------------------------------------------
# ExRK: the Butcher coefficients give 1st..4th order (gtest_rk.cpp logic)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rk_convergence_order(order):
    g = oracle.Grid.box(2, 4, [2, 2])
    x, y = g.get("x"), g.get("y")
    u0 = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    T = 0.02

    def run(nsteps):
        st = [u0.copy(), u0.copy()]
        for s in range(nsteps):
            g.burgers_step(st, 1.0, T / nsteps, mode=2, order=order)
        return st[0]

    ref = run(400)
    e1 = np.abs(run(8) - ref).max()
    e2 = np.abs(run(16) - ref).max()
    rate = np.log2(e1 / e2)
    assert abs(rate - order) < 0.35, rate


def test_heat_mode_evolves_first_component_only():
    g = oracle.Grid.box(2, 3, [2, 2])
    x = g.get("x")
    st = [np.sin(2 * np.pi * x), np.cos(2 * np.pi * x)]
    c1 = st[1].copy()
    g.burgers_step(st, 0.1, 1e-3, mode=2)
    assert np.array_equal(st[1], c1)


def test_burgers_rhs_closed_form():
    """The nonlinear, pure-advection and heat right-hand sides of
    or_burgers_rhs, GGFX-smoothed as the stepper does, against the strong
    form at the nodes for a smooth periodic field (spectral accuracy at p=10):
    nonlinear dudt_c = -(u.grad) u_c + nu Lap u_c; advection dudt = -c.grad s
    + nu Lap s; heat dudt_c = nu Lap u_c.  A dropped term, a wrong sign or a
    transposed operand fails by O(1)."""
    og = oracle.Grid.box(2, 10, [4, 4])
    x, y = og.get("x"), og.get("y")
    k, A, B, nu = 2 * np.pi, 0.7, 0.4, 0.03
    sx, cx, sy, cy = np.sin(k * x), np.cos(k * x), np.sin(k * y), np.cos(k * y)
    u0, u1 = A * sx * cy, B * cx * sy
    u0x, u0y = A * k * cx * cy, -A * k * sx * sy
    u1x, u1y = -B * k * sx * sy, B * k * cx * cy
    want = [-(u0 * u0x + u1 * u0y) - nu * 2 * k * k * u0,
            -(u0 * u1x + u1 * u1y) - nu * 2 * k * k * u1]
    got = og.burgers_rhs([u0, u1], nu, mode=0)
    scale = max(np.abs(w).max() for w in want)
    for c in range(2):
        err = np.abs(og.dss(got[c], "smooth") - want[c]).max() / scale
        assert err <= 1e-7, (c, err)
    # pure advection of s by the velocity (u0, u1)
    s = 0.5 * cx * cy
    sxd, syd = -0.5 * k * sx * cy, -0.5 * k * cx * sy
    got = og.burgers_rhs([s, u0, u1], nu, mode=1)
    want_s = -(u0 * sxd + u1 * syd) - nu * 2 * k * k * s
    assert np.abs(og.dss(got[0], "smooth") - want_s).max() <= 1e-7 * np.abs(want_s).max()
    # heat
    got = og.burgers_rhs([u0, u1], nu, mode=2)
    assert np.abs(og.dss(got[0], "smooth") + nu * 2 * k * k * u0).max() <= \
        1e-7 * nu * 2 * k * k * A


def _diag_grid(kind):
    if kind == "box2_dirichlet":
        return oracle.Grid.box(2, 3, [2, 3], bdy=[6] * 4)
    if kind == "box2_periodic":
        return oracle.Grid.box(2, 4, [2, 2], bdy=[0] * 4)
    if kind == "box3_periodic":
        return oracle.Grid.box(3, 2, [2, 2, 2], bdy=[0] * 6)
    if kind == "terrain2":
        return oracle.Grid.terrain(3, [2, 3], 0.3)
    if kind == "terrain3":
        return oracle.Grid.terrain3(2, [2, 2, 2], 0.3)
    return oracle.Grid.icos(3, 1, radius=1.3)


@pytest.mark.parametrize("pfield", [False, True], ids=["p_scalar", "p_field"])
@pytest.mark.parametrize("kind", ["box2_dirichlet", "box2_periodic", "box3_periodic", "terrain2",
                                  "terrain3", "sphere"])
def test_jacobi_diagonal_by_unit_vectors(kind, pfield):
    """or_helmholtz_diag_inv (the Jacobi preconditioner, no reference counterpart) against the
    assembled diagonal found by brute force: apply (p L + q M) to the
    continuous unit field of each node group, assemble (GGFX Sum) and read
    the group's own entry.  On the terrain and sphere meshes the off-diagonal
    metric terms reach the diagonal at element edges (D(n_a, n_a) is non-zero
    there); dropping or halving them fails the check."""
    og = _diag_grid(kind)
    q = 1.0 + 0.5 * np.random.default_rng(3).random(og.ndof)
    q = og.dss(q, "smooth")
    # p as a full field varies inside the elements, so it must enter the
    # diagonal at every node of the line sum (the reference multiplies it in
    # pointwise before D^T); a continuous field, as a Laplacian factor is
    p = og.dss(0.2 + 0.2 * np.random.default_rng(5).random(og.ndof), "smooth") if pfield else 0.3
    dinv = og.helmholtz_diag_inv(p, q)
    offs, mem = og.ggfx()
    done = np.zeros(og.ndof, dtype=bool)
    for l in range(og.ndof):
        if done[l]:
            continue
        group = mem[offs[l]:offs[l + 1]]
        e = np.zeros(og.ndof)
        e[group] = 1.0
        a = og.dss(og.helmholtz(e, p, q), "sum")
        for g in group:
            assert abs(dinv[g] * a[g] - 1.0) <= 1e-12, (l, g)
        done[group] = True


# ---------------------------------------------------------------------------
# Sphere path (GE_2DEMBEDDED): constrain2sphere, the embedded gradient and
# GAdvect against closed forms of the sphere's geometry
# ---------------------------------------------------------------------------
def _sphere_xyz(og):
    return [og.get(c) for c in "xyz"]


def test_constrain2sphere_tangent_fields_unchanged():
    """A tangent field (x . v = 0) is a fixed point of the projection, also
    with the reference's in-place aliasing (gmtk.ipp:2972-2974): v0' = v0
    exactly up to rounding, so v1 and v2 see the original v0."""
    og = oracle.Grid.icos(4, 1, radius=1.7)
    X = _sphere_xyz(og)
    rng = np.random.default_rng(11)
    a = rng.standard_normal(3)
    # v = a x X is tangent everywhere
    v = [a[1] * X[2] - a[2] * X[1], a[2] * X[0] - a[0] * X[2], a[0] * X[1] - a[1] * X[0]]
    got = og.constrain2sphere([c.copy() for c in v])
    scale = max(np.abs(c).max() for c in v)
    for c in range(3):
        assert np.abs(got[c] - v[c]).max() <= 1e-14 * scale, c


def test_constrain2sphere_radial_field_aliased_closed_form():
    """A radial field v = X: the true projection is 0, but the reference
    updates in place, so v1 and v2 are formed from the already-projected v0.
    With r^2 = |X|^2 (every node on the sphere): v0' = X(r^2 - |X|^2)/r^2 = 0,
    v1' = y x^2 / r^2 (the -v0' y x term vanishes, the rest is y(r^2 - y^2 - z^2)),
    v2' = z (x^2 (r^2 - y^2) / r^2 + y^2) / r^2.  A projection without the
    aliasing (all three from the original v) would give 0 in every component."""
    R = 1.3
    og = oracle.Grid.icos(3, 1, radius=R)
    X = _sphere_xyz(og)
    x, y, z = X
    got = og.constrain2sphere([c.copy() for c in X])
    r2 = R * R
    want = [np.zeros_like(x), y * x * x / r2, z * (x * x * (r2 - y * y) / r2 + y * y) / r2]
    for c in range(3):
        assert np.abs(got[c] - want[c]).max() <= 1e-13 * R, c
    assert np.abs(got[1]).max() > 0.1  # the aliasing is visible, not round-off


@pytest.mark.parametrize("p", [6, 8, 10])
def test_embedded_gradient_of_coordinates_is_tangent_projection(p):
    """GGrid::deriv on a GE_2DEMBEDDED grid (ggrid.ipp:1163-1191): the surface
    gradient of the coordinate function x_k is the tangent projection of e_k,
    d x_k / d x_j = delta_kj - x_k x_j / R^2.  The metric comes from the
    embedded shape functions, the derivative from D on nodal values, so the
    two meet to spectral accuracy: the error falls ~100x per +2 in p (1e-9 at
    p=8, 1e-11 at p=10 on the level-1 sphere).  A transposed metric, a
    dropped xi-direction or a wrong radial column fails at O(1)."""
    og = oracle.Grid.icos(p, 1, radius=1.0)
    X = _sphere_xyz(og)
    R2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    err = 0.0
    for k in range(3):
        for j in range(3):
            d = og.deriv(X[k].copy(), j + 1)
            want = (1.0 if k == j else 0.0) - X[k] * X[j] / R2
            err = max(err, np.abs(d - want).max())
    bound = {6: 1e-6, 8: 1e-8, 10: 1e-10}[p]
    assert err <= bound, err


@pytest.mark.parametrize("p", [6, 8, 10])
def test_embedded_advection_solid_body_rotation_closed_forms(p):
    """GAdvect::apply (gadvect.ipp:57-99) on the sphere: po = M (c . grad z).
    Rotation about z (c = (-y, x, 0)) moves z along its level sets, so
    c . grad z = 0; rotation about x (c = (0, -z, y)) gives c . grad z = y."""
    og = oracle.Grid.icos(p, 1, radius=1.0)
    x, y, z = _sphere_xyz(og)
    mass = og.get("mass")
    zero = np.zeros_like(x)
    po = og.advect(z.copy(), [-y, x, zero])
    bound = {6: 1e-6, 8: 1e-8, 10: 1e-10}[p]
    assert np.abs(po / mass).max() <= bound
    po = og.advect(z.copy(), [zero, -z, y])
    assert np.abs(po / mass - y).max() <= bound


def _assembled_advection_matrix(og, c):
    """The pure-advection RHS K = Smooth(M^-1 (-GAdvect(u, c))) of GBurgers
    (ν = 0) as a matrix on the unique nodes (one column per node group)."""
    offs, mem = og.ggfx()
    n = og.ndof
    uid = -np.ones(n, dtype=np.int64)
    nu = 0
    for g in range(len(offs) - 1):
        uid[mem[offs[g]:offs[g + 1]]] = nu
        nu += 1
    for i in range(n):
        if uid[i] < 0:
            uid[i] = nu
            nu += 1
    rep = np.zeros(nu, dtype=np.int64)
    rep[uid] = np.arange(n)
    A = np.zeros((nu, nu))
    for j in range(nu):
        e = (uid == j).astype(float)
        k = og.dss(og.imass_apply(-og.advect(e, c)), "smooth")
        A[:, j] = k[rep]
    return A


def test_advection_spectrum_flat_box_neutral_sphere_growing():
    """Why pure advection (ν = 0) on the icosahedral sphere grows without
    bound: the collocation form of GAdvect on curved (embedded) elements has
    eigenvalues with positive real part, and they grow with resolution.  On
    a flat periodic box the same operator with a variable divergence-free
    swirl is neutral to round-off (max Re λ ~ 1e-13 against |Im λ| ~ 250);
    on the sphere under solid-body rotation (|c| ≤ 2π) max Re λ is 0.5 at
    level 0, p=3, and rises with p.  The operator itself is pinned above
    (gradient and advection closed forms), so the growth is the scheme's:
    grid-scale modes of the under-integrated, non-skew advection on curved
    elements, which the reference damps with filters (not on this path)."""
    og = oracle.Grid.box(2, 5, [4, 4])
    x, y = og.get("x"), og.get("y")
    tp = 2 * np.pi
    c = [tp * np.sin(tp * x) * np.cos(tp * y), -tp * np.cos(tp * x) * np.sin(tp * y)]
    lam = np.linalg.eigvals(_assembled_advection_matrix(og, c))
    assert np.abs(lam.imag).max() > 100
    assert lam.real.max() < 1e-10 * np.abs(lam.imag).max()
    res = []
    for p in (3, 5):
        og = oracle.Grid.icos(p, 0, radius=1.0)
        X = _sphere_xyz(og)
        ax = np.array([-np.sin(np.pi / 4), 0.0, np.cos(np.pi / 4)]) * tp
        c = [ax[1] * X[2] - ax[2] * X[1], ax[2] * X[0] - ax[0] * X[2],
             ax[0] * X[1] - ax[1] * X[0]]
        lam = np.linalg.eigvals(_assembled_advection_matrix(og, c))
        res.append(lam.real.max())
    assert res[0] > 0.3 and res[1] > 1.3 * res[0], res


# ---------------------------------------------------------------------------
# Quadratic dealiasing (GGrid::dealias, ggrid.ipp:2351-2503)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim,p,pqs", [(2, 4, (4, 6, 8)), (3, 3, (3, 5))])
def test_dealias_identities(dim, p, pqs):
    """dealias(1, 1) = 1 for every pq (the GLL weights are the exact
    integrals of the Lagrange basis); with pq = p the interpolation is the
    identity and dealias is the pointwise product."""
    og = oracle.Grid.box(dim, p, [3] * dim)
    one = np.ones(og.ndof)
    for pq in pqs:
        assert np.abs(og.dealias(one, one, pq) - 1.0).max() <= 1e-14
    x, y = og.get("x"), og.get("y")
    u, v = np.sin(3 * x) + y, np.cos(2 * y) * x
    assert np.abs(og.dealias(u, v, p) - u * v).max() <= 1e-14


def test_dealias_is_the_exact_galerkin_projection():
    """For a product f = v1 v2 of degree <= p on each element and pq high
    enough that the order-pq GLL quadrature integrates phi_i f exactly,
    dealias returns (M f)_i / W_i per direction: the exact 1-D mass matrix
    M_im = int phi_i phi_m, formed here independently with numpy's Lagrange
    interpolation and 20-point Gauss-Legendre quadrature."""
    p, pq = 4, 6
    og = oracle.Grid.box(2, p, [1, 1], bdy=[0] * 4)
    xi, w, _ = oracle.gll_basis(p)
    gx, gw = np.polynomial.legendre.leggauss(20)
    N = p + 1
    phi = np.zeros((N, gx.size))
    for i in range(N):
        others = [xi[m] for m in range(N) if m != i]
        phi[i] = np.prod([(gx - o) / (xi[i] - o) for o in others], axis=0)
    M = (phi * gw) @ phi.T
    x, y = og.get("x"), og.get("y")
    v1, v2 = x, y * y
    got = og.dealias(v1, v2, pq).reshape(N, N)  # [j, i]
    fx = (xi + 1) / 2              # x on the element's nodes
    fy = ((xi + 1) / 2) ** 2       # y^2 on the element's nodes
    want = np.outer(M @ fy / w, M @ fx / w)
    assert np.abs(got - want).max() <= 1e-13
