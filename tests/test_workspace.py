"""Mesh workspace: gf_mesh_reserve allocates the scratch of the selected
entry points ahead of the calls, and GF_RESERVE_FREEZE turns any later
allocation into GF_ERR_ALLOC, so a frozen mesh proves that CG solves (FP64
and mixed), RK steps and the checkpointed adjoint allocate nothing inside
the call.  Results on a frozen mesh equal those of an ordinary mesh bit for
bit (same kernels, same buffers' contents).  Also the Jacobi preconditioner's
positivity check (GF_ERR_PRECOND for a diagonal that is not positive)."""
import numpy as np
import pytest
import torch

import geoflow_amd as ga
from geoflow_amd import _capi, workloads
from geoflow_amd.device import Mesh

pytestmark = pytest.mark.gpu

R = _capi


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_frozen_mesh_without_reservation_refuses_to_allocate(ctx):
    print("workspace: frozen without reservation", flush=True)
    bg = ga.BoxGrid(3, 4, [3, 2, 2], bdy=[0] * 6)
    m = Mesh(ctx, bg)
    m.reserve(R.GF_RESERVE_FREEZE)
    b = dev(workloads.unique_normal(bg, seed=3))
    with pytest.raises(_capi.GeoflowError) as ei:
        m.cg_solve(b, p=0.1, q=1.0, maxit=20, tol=1e-8, precond=1)
    assert ei.value.code == R.GF_ERR_ALLOC


@pytest.mark.parametrize("bdy", [[0] * 6, [6] * 6], ids=["dirichlet", "periodic"])
def test_reserved_frozen_mesh_runs_cg_rk_adjoint_without_allocating(ctx, bdy):
    print("workspace: reserved and frozen", bdy[0], flush=True)
    bg = ga.BoxGrid(3, 5, [3, 3, 2], bdy=bdy)
    plain = Mesh(ctx, bg)
    frozen = Mesh(ctx, bg)
    frozen.reserve(R.GF_RESERVE_RK | R.GF_RESERVE_CG | R.GF_RESERVE_CG_MIXED |
                   R.GF_RESERVE_ADJOINT | R.GF_RESERVE_FREEZE,
                   nstate=3, cg_maxit=300, adj_nsteps=7, adj_ckpt=3, adj_order=4)
    free0 = torch.cuda.mem_get_info()[0]
    b = dev(workloads.unique_normal(bg, seed=4))
    for prec in (0, 1):
        xa, ia = plain.cg_solve(b, p=0.1, q=1.0, maxit=300, tol=1e-10, normtype=4,
                                precond=1, precision=prec)
        xb, ib = frozen.cg_solve(b, p=0.1, q=1.0, maxit=300, tol=1e-10, normtype=4,
                                 precond=1, precision=prec)
        assert ia["status"] == ib["status"] == 0
        assert ia["iters"] == ib["iters"]
        assert np.array_equal(host(xa), host(xb)), prec
    if bdy[0] == 6:
        comps = [0.5 * workloads.unique_normal(bg, seed=20 + c) for c in range(3)]
        ua = [dev(c) for c in comps]
        ub = [dev(c) for c in comps]
        plain.burgers_steps(ua, 0.02, 1e-3, 3)
        frozen.burgers_steps(ub, 0.02, 1e-3, 3)
        plain.burgers_step(ua, 0.02, 1e-3)
        frozen.burgers_step(ub, 0.02, 1e-3)
        for c in range(3):
            assert np.array_equal(host(ua[c]), host(ub[c])), c
        u0 = [dev(c) for c in comps]
        la = [dev(np.ones(bg.ndof) * (c + 1)) for c in range(3)]
        lb = [dev(np.ones(bg.ndof) * (c + 1)) for c in range(3)]
        plain.burgers_adjoint(u0, la, 0.02, 1e-3, 7, ckpt=3)
        frozen.burgers_adjoint(u0, lb, 0.02, 1e-3, 7, ckpt=3)
        for c in range(3):
            assert np.array_equal(host(la[c]), host(lb[c])), c
    torch.cuda.synchronize()
    # the frozen mesh allocated nothing during the calls (the plain mesh did;
    # device memory can only have shrunk by the plain mesh's workspace)
    assert torch.cuda.mem_get_info()[0] <= free0


def test_jacobi_diagonal_must_be_positive(ctx):
    print("workspace: jacobi positivity", flush=True)
    bg = ga.BoxGrid(3, 3, [2, 2, 2], bdy=[0] * 6)
    m = Mesh(ctx, bg)
    m.helmholtz_diag_inv(0.1, 1.0)  # positive: fine
    with pytest.raises(_capi.GeoflowError) as ei:
        m.helmholtz_diag_inv(0.0, 0.0)
    assert ei.value.code == R.GF_ERR_PRECOND
    with pytest.raises(_capi.GeoflowError) as ei:
        m.helmholtz_diag_inv(-1.0, 0.0)
    assert ei.value.code == R.GF_ERR_PRECOND
