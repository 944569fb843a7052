"""Adjoint of the Burgers RHS and RK step, and the checkpointed gradient.

The adjoint has no counterpart in the reference; what pins it is the forward
map it differentiates (GBurgers::dudt_impl, gburgers.ipp:203-257, and
GExRKStepper::step_b, gexrk_stepper.ipp:252-341, both already pinned to the
oracle and the GPU path at 1e-12):

CPU (oracle):
  * the nonlinear RHS is quadratic in u, so a central difference of it is
    exact up to round-off: <w, (F(u+ev) - F(u-ev)) / 2e> = <(dF/du)^T w, v>
    to ~1e-12 for any e (a dropped or mis-signed term fails at O(1));
  * one RK step (orders 1-4) and a chain of steps against central
    differences of the oracle's forward step (e = 1e-6; truncation O(e^2)).
GPU:
  * gf_burgers_rhs_adjoint and gf_burgers_step_adjoint against the oracle at
    1e-12 (2-D, 3-D general path, 3-D structured-box path);
  * gf_burgers_adjoint (checkpointed, uneven last segment) against the oracle
    gradient chained step by step, and its u(T) against the forward steps;
  * the workload of the adjoint benchmark (3-D periodic box, 16^3 elements,
    p=6, nu=0.02, 200 RK4 steps, dt=5e-5, checkpoints every 20 steps, loss
    0.5 ||u(T) - u_target||_M^2): gradient against central differences
    (e = 1e-6) along 8 random directions at 1e-6 relative, through a
    torch.autograd.Function whose backward calls the C ABI adjoint.
"""
import numpy as np
import pytest

import oracle
from conftest import relerr
from geoflow_amd.workloads import fourier_field_3d


def _field(og, seed, kmax=2):
    x, y = og.get("x"), og.get("y")
    if og.dim == 2:
        return fourier_field_3d(x, y, 0 * x, seed=seed, kmax=kmax, ncomp=2)
    return fourier_field_3d(x, y, og.get("z"), seed=seed, kmax=kmax, ncomp=3)


def _dot(a, b):
    return float(sum(np.dot(x, y) for x, y in zip(a, b)))


GRIDS = [
    dict(dim=2, p=4, ne=[3, 3]),
    dict(dim=3, p=3, ne=[2, 2, 2]),
]


def _gid(c):
    return "d%d_p%d" % (c["dim"], c["p"])


@pytest.mark.parametrize("c", GRIDS, ids=_gid)
def test_oracle_rhs_adjoint_is_exact_transpose(c):
    og = oracle.Grid.box(c["dim"], c["p"], c["ne"])
    rng = np.random.default_rng(5)
    u = [1.3 * f for f in _field(og, 1)]
    w = [rng.standard_normal(og.ndof) for _ in u]
    v = [rng.standard_normal(og.ndof) for _ in u]
    nu, e = 0.05, 1e-2
    Fp = og.burgers_rhs([a + e * b for a, b in zip(u, v)], nu)
    Fm = og.burgers_rhs([a - e * b for a, b in zip(u, v)], nu)
    lhs = _dot(w, [(p - m) / (2 * e) for p, m in zip(Fp, Fm)])
    lam = og.burgers_rhs_adjoint(u, w, [np.zeros(og.ndof) for _ in u], nu)
    rhs = _dot(lam, v)
    assert abs(lhs - rhs) <= 1e-11 * abs(rhs)
    # accumulation into lam (+=)
    lam2 = og.burgers_rhs_adjoint(u, w, [np.ones(og.ndof) for _ in u], nu)
    for a, b in zip(lam2, lam):
        assert np.allclose(a - 1.0, b, rtol=0, atol=1e-13 * np.abs(b).max())


@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("c", GRIDS, ids=_gid)
def test_oracle_step_adjoint_central_difference(c, order):
    og = oracle.Grid.box(c["dim"], c["p"], c["ne"])
    rng = np.random.default_rng(7)
    u = _field(og, 2)
    w = [rng.standard_normal(og.ndof) for _ in u]
    nu, dt, e = 0.05, 2e-3, 1e-6

    def J(st):
        st = [a.copy() for a in st]
        og.burgers_step(st, nu, dt, order=order)
        return _dot(w, st)

    lam = og.burgers_step_adjoint(u, [a.copy() for a in w], nu, dt, order=order)
    for k in range(3):
        v = [rng.standard_normal(og.ndof) for _ in u]
        fd = (J([a + e * b for a, b in zip(u, v)]) - J([a - e * b for a, b in zip(u, v)])) / (2 * e)
        ad = _dot(lam, v)
        assert abs(fd - ad) <= 1e-7 * abs(ad), (k, fd, ad)


def _oracle_gradient(og, u0, target, mass, nu, dt, nsteps, order=4):
    """dJ/du(0) for J = 0.5 sum mass (u(T) - target)^2, chaining the oracle's
    step adjoint backwards over stored states."""
    states = [[a.copy() for a in u0]]
    st = [a.copy() for a in u0]
    for s in range(nsteps):
        og.burgers_step(st, nu, dt, order=order)
        states.append([a.copy() for a in st])
    lam = [mass * (a - t) for a, t in zip(st, target)]
    for s in range(nsteps - 1, -1, -1):
        og.burgers_step_adjoint(states[s], lam, nu, dt, order=order)
    return lam, st


def test_oracle_gradient_chain_central_difference():
    og = oracle.Grid.box(3, 3, [2, 2, 2])
    u0, tg = _field(og, 3), _field(og, 4)
    mass = og.get("mass")
    nu, dt, n = 0.02, 2e-3, 6

    def J(u):
        st = [a.copy() for a in u]
        for s in range(n):
            og.burgers_step(st, nu, dt)
        return 0.5 * sum(float(np.dot(mass, (a - t) ** 2)) for a, t in zip(st, tg))

    g, _ = _oracle_gradient(og, u0, tg, mass, nu, dt, n)
    rng = np.random.default_rng(11)
    for k in range(3):
        v = [rng.standard_normal(og.ndof) for _ in u0]
        e = 1e-6
        fd = (J([a + e * b for a, b in zip(u0, v)]) - J([a - e * b for a, b in zip(u0, v)])) / (2 * e)
        ad = _dot(g, v)
        assert abs(fd - ad) <= 1e-7 * abs(ad), (k, fd, ad)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
TOL = 1e-12

GPU_GRIDS = [
    dict(dim=2, p=4, ne=[4, 3], structured=False),
    dict(dim=3, p=4, ne=[3, 2, 2], structured=False),
    dict(dim=3, p=5, ne=[2, 3, 2], structured=True),
]


def _gpu_gid(c):
    return "d%d_p%d_%s" % (c["dim"], c["p"], "box" if c["structured"] else "general")


def _setup(ctx, c):
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    og = oracle.Grid.box(c["dim"], c["p"], c["ne"])
    bg = ga.BoxGrid(c["dim"], c["p"], c["ne"])
    return og, Mesh(ctx, bg, structured=c["structured"])


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("c", GPU_GRIDS, ids=_gpu_gid)
def test_rhs_adjoint_matches_oracle(ctx, c):
    print("rhs_adjoint", c, flush=True)
    og, m = _setup(ctx, c)
    rng = np.random.default_rng(21)
    u = _field(og, 1)
    w = [rng.standard_normal(og.ndof) for _ in u]
    l0 = [rng.standard_normal(og.ndof) for _ in u]
    ref = og.burgers_rhs_adjoint(u, w, [a.copy() for a in l0], 0.03)
    got = m.burgers_rhs_adjoint([_dev(a) for a in u], [_dev(a) for a in w],
                                [_dev(a) for a in l0], 0.03)
    for k in range(len(u)):
        assert relerr(_host(got[k]), ref[k]) <= TOL, k


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("c", GPU_GRIDS, ids=_gpu_gid)
def test_step_adjoint_matches_oracle(ctx, c, order):
    print("step_adjoint", c, order, flush=True)
    og, m = _setup(ctx, c)
    rng = np.random.default_rng(22)
    u = _field(og, 2)
    w = [rng.standard_normal(og.ndof) for _ in u]
    ref = og.burgers_step_adjoint(u, [a.copy() for a in w], 0.04, 1e-3, order=order)
    got = m.burgers_step_adjoint([_dev(a) for a in u], [_dev(a) for a in w], 0.04, 1e-3,
                                 order=order)
    for k in range(len(u)):
        assert relerr(_host(got[k]), ref[k]) <= TOL, k


@pytest.mark.gpu
@pytest.mark.parametrize("c", GPU_GRIDS[1:], ids=_gpu_gid)
def test_checkpointed_gradient_matches_oracle(ctx, c):
    print("checkpointed gradient", c, flush=True)
    og, m = _setup(ctx, c)
    u0, tg = _field(og, 3), _field(og, 4)
    mass = og.get("mass")
    nu, dt, n = 0.02, 1e-3, 13
    g_ref, uT_ref = _oracle_gradient(og, u0, tg, mass, nu, dt, n)
    st = [_dev(a) for a in u0]
    uT = [m.new() for _ in u0]
    # lam = dJ/du(T) needs u(T): a forward pass first, as a caller would
    fw = [a.clone() for a in st]
    m.burgers_steps(fw, nu, dt, n)
    lam = [_dev(mass) * (a - _dev(t)) for a, t in zip(fw, tg)]
    m.burgers_adjoint(st, lam, nu, dt, n, ckpt=5, uT=uT)
    for k in range(3):
        assert relerr(_host(uT[k]), uT_ref[k]) <= TOL
        assert relerr(_host(fw[k]), uT_ref[k]) <= TOL
        assert relerr(_host(lam[k]), g_ref[k]) <= 1e-11, k
        assert np.array_equal(_host(st[k]), u0[k])  # input state untouched


def _rollout_function():
    import torch

    class BurgersRollout(torch.autograd.Function):
        """u(T) from u(0) by nsteps RK steps; backward = checkpointed adjoint."""

        @staticmethod
        def forward(fctx, mesh, nu, dt, nsteps, ckpt, *u0):
            fctx.mesh, fctx.prm = mesh, (nu, dt, nsteps, ckpt)
            fctx.save_for_backward(*u0)
            uT = [a.detach().clone() for a in u0]
            mesh.burgers_steps(uT, nu, dt, nsteps)
            return tuple(uT)

        @staticmethod
        def backward(fctx, *gout):
            nu, dt, nsteps, ckpt = fctx.prm
            u0 = [a.detach().contiguous() for a in fctx.saved_tensors]
            lam = [g.detach().clone().contiguous() for g in gout]
            fctx.mesh.burgers_adjoint(u0, lam, nu, dt, nsteps, ckpt=ckpt)
            return (None, None, None, None, None) + tuple(lam)

    return BurgersRollout


@pytest.mark.gpu
def test_adjoint_gradient_fd_16cube_p6_200_steps(ctx):
    """16^3 elements, p=6 (1.4M DOF), nu=0.02, 200 RK4 steps at dt=5e-5,
    checkpoints every 20 steps; 8 random directions, e = 1e-6."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    print("adjoint fd 16^3 p=6", flush=True)
    bg = ga.BoxGrid(3, 6, [16, 16, 16])
    m = Mesh(ctx, bg)
    X = [torch.from_numpy(a).cuda() for a in bg.xyz]
    u0 = [a.contiguous() for a in fourier_field_3d(*X, seed=101, xp=torch)]
    tg = [a.contiguous() for a in fourier_field_3d(*X, seed=202, xp=torch)]
    mass = torch.from_numpy(m.get("mass")).cuda()
    nu, dt, n, ck = 0.02, 5e-5, 200, 20
    Roll = _rollout_function()

    def loss_of(uT):
        return 0.5 * sum((mass * (a - t) ** 2).sum() for a, t in zip(uT, tg))

    leaf = [a.clone().requires_grad_(True) for a in u0]
    L = loss_of(Roll.apply(m, nu, dt, n, ck, *leaf))
    L.backward()
    grad = [a.grad.detach() for a in leaf]
    # the same gradient without autograd
    fw = [a.clone() for a in u0]
    m.burgers_steps(fw, nu, dt, n)
    lam = [mass * (a - t) for a, t in zip(fw, tg)]
    m.burgers_adjoint([a.clone() for a in u0], lam, nu, dt, n, ckpt=ck)
    for a, b in zip(grad, lam):
        assert float((a - b).abs().max()) <= 1e-14 * float(b.abs().max())

    def J(u):
        st = [a.clone() for a in u]
        m.burgers_steps(st, nu, dt, n)
        return float(loss_of(st))

    gen = torch.Generator(device="cpu").manual_seed(303)
    worst = 0.0
    for k in range(8):
        v = [torch.randn(m.ndof, generator=gen, dtype=torch.float64).cuda() for _ in u0]
        v = [m.dss(a, "smooth") for a in v]  # continuous directions
        e = 1e-6
        fd = (J([a + e * b for a, b in zip(u0, v)]) - J([a - e * b for a, b in zip(u0, v)])) / (2 * e)
        ad = float(sum((a * b).sum() for a, b in zip(grad, v)))
        rel = abs(fd - ad) / abs(ad)
        print("direction %d: fd %.12e adjoint %.12e rel %.2e" % (k, fd, ad, rel), flush=True)
        worst = max(worst, rel)
    assert worst <= 1e-6


@pytest.mark.gpu
def test_gio_checkpointed_gradient_equals_device_checkpoints(ctx, tmp_path):
    """gf_burgers_adjoint_gio keeps its checkpoints as GIO state files: the
    gradient and u(T) equal gf_burgers_adjoint's bit for bit (7 steps,
    checkpoints every 3, so the last segment is short), and checkpoint c is
    a GeoFLOW state file of the state at step 3c (header cycle/time/orders,
    one field per component) equal to the forward steps bit for bit."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd import gio
    from geoflow_amd.device import Mesh
    print("adjoint gio checkpoints", flush=True)
    bg = ga.BoxGrid(3, 4, [3, 2, 2])
    m = Mesh(ctx, bg)
    u0 = [torch.from_numpy(a).cuda() for a in fourier_field_3d(*bg.xyz, seed=3, kmax=2)]
    lam0 = [torch.from_numpy(a).cuda() for a in fourier_field_3d(*bg.xyz, seed=4, kmax=2)]
    nu, dt, n, ck = 0.02, 1e-3, 7, 3
    la = [a.clone() for a in lam0]
    lb = [a.clone() for a in lam0]
    ua = [torch.empty_like(a) for a in u0]
    ub = [torch.empty_like(a) for a in u0]
    m.burgers_adjoint(u0, la, nu, dt, n, ckpt=ck, uT=ua)
    prefix = tmp_path / "ck"
    m.burgers_adjoint(u0, lb, nu, dt, n, ckpt=ck, uT=ub, gio_prefix=prefix)
    for a, b in zip(la + ua, lb + ub):
        assert torch.equal(a, b)
    st = [a.clone() for a in u0]
    for c in range(3):
        meta, rows, keys, fields = gio.read_state(str(prefix) + ".%d.gio" % c, nfields=3)
        assert meta["cycle"] == ck * c and abs(meta["time"] - ck * c * dt) < 1e-15
        assert meta["multivar"] == 1 and meta["nelems"] == bg.nelems
        assert list(rows[0]) == [4, 4, 4]
        assert np.array_equal(keys, np.arange(bg.nelems))
        for f, s in zip(fields, st):
            assert np.array_equal(f, s.cpu().numpy())
        m.burgers_steps(st, nu, dt, ck)
