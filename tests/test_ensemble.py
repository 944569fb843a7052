"""Batched ensembles (geoflow_amd/ensemble.py): members of uneven element
counts and orders, members of one order concatenated into one mesh with
block-diagonal GGFX groups.

CPU: the concatenated GGFX maps are exactly the members' maps shifted (no
group crosses a member).
GPU: each member's implicit-viscous step computed in the batch equals the
same step on the member's own mesh; FP32-inner solves agree with FP64 to
the CG tolerance; the batched RK4 adjoint gradient of every member equals
the member's own gradient.
"""
import numpy as np
import pytest

import geoflow_amd as ga
from geoflow_amd.ensemble import MemberSpec, _UnionGrid, draw_members
from geoflow_amd.workloads import fourier_field_3d
from conftest import relerr

MEMBERS = [MemberSpec(3, 2), MemberSpec(4, 3), MemberSpec(3, 3), MemberSpec(4, 2),
           MemberSpec(5, 2)]


def test_union_ggfx_is_block_diagonal():
    grids = [ga.BoxGrid(3, 3, [ne] * 3) for ne in (2, 3, 2)]
    ug = _UnionGrid(grids)
    assert ug.ndof == sum(g.ndof for g in grids)
    off = 0
    for g in grids:
        n = g.ndof
        o = ug.ggfx_offs[off:off + n + 1]
        assert np.array_equal(o - o[0], g.ggfx_offs)
        mem = ug.ggfx_members[o[0]:o[-1]]
        assert np.array_equal(mem - off, g.ggfx_members)
        off += n
    assert ug.ggfx_offs[-1] == ug.ggfx_members.size


def test_draw_members_by_seed():
    a, b = draw_members(48, seed=7), draw_members(48, seed=7)
    assert [(m.p, m.ne) for m in a] == [(m.p, m.ne) for m in b]
    assert {m.p for m in a} <= {4, 6, 8} and {m.ne for m in a} <= {8, 12, 16, 24}


def _ic(seed):
    def fill(i, g):
        base = fourier_field_3d(*g.xyz, seed=seed, kmax=2)
        rng = np.random.default_rng(seed + 1000 + i)
        return [b + 1e-3 * rng.standard_normal(g.ndof) for b in base]
    return fill


@pytest.mark.gpu
def test_batched_imex_step_matches_members_alone(ctx):
    from geoflow_amd.device import Mesh
    from geoflow_amd.ensemble import Ensemble
    import torch
    print("ensemble imex", flush=True)
    ens = Ensemble(ctx, MEMBERS)
    assert sorted(ens.mesh) == [3, 4, 5]
    st = ens.new_state(_ic(5))
    st32 = {p: [a.clone() for a in v] for p, v in st.items()}
    nu, dt = 0.05, 2e-3
    ens.imex_step(st, nu, dt, precision=0, tol=1e-13)
    ens.imex_step(st32, nu, dt, precision=1, tol=1e-13)
    for i, ms in enumerate(MEMBERS):
        g = ga.BoxGrid(3, ms.p, [ms.ne] * 3)
        m = Mesh(ctx, g, structured=False)
        u0 = _ic(5)(i, g)
        us = [torch.from_numpy(a).cuda() for a in u0]
        new = []
        for c in range(3):
            b = m.mass_apply(us[c]) - dt * m.advect(us[c], us)
            x, info = m.cg_solve(b, x=us[c].clone(), p=nu * dt, q=1.0, maxit=2000, tol=1e-13,
                                 normtype=4, precond=1)
            assert info["status"] == 0
            new.append(x.cpu().numpy())
        got = ens.member(st, i)
        got32 = ens.member(st32, i)
        for c in range(3):
            assert relerr(got[c], new[c]) <= 1e-10, (i, c)
            assert relerr(got32[c], new[c]) <= 1e-10, (i, c)


@pytest.mark.gpu
def test_batched_rk4_gradient_matches_members_alone(ctx):
    from geoflow_amd.device import Mesh
    from geoflow_amd.ensemble import Ensemble
    import torch
    print("ensemble gradient", flush=True)
    ens = Ensemble(ctx, MEMBERS)
    st0, tg = ens.new_state(_ic(8)), ens.new_state(_ic(9))
    nu, dt, n = 0.05, 1e-3, 6
    grads, _ = ens.rk4_gradient(st0, tg, nu, dt, n, ckpt=4)
    for i, ms in enumerate(MEMBERS):
        g = ga.BoxGrid(3, ms.p, [ms.ne] * 3)
        m = Mesh(ctx, g, structured=False)
        u0 = [torch.from_numpy(a).cuda() for a in _ic(8)(i, g)]
        t = [torch.from_numpy(a).cuda() for a in _ic(9)(i, g)]
        fw = [a.clone() for a in u0]
        m.burgers_steps(fw, nu, dt, n)
        mass = torch.from_numpy(m.get("mass")).cuda()
        lam = [mass * (a - b) for a, b in zip(fw, t)]
        m.burgers_adjoint([a.clone() for a in u0], lam, nu, dt, n, ckpt=4)
        p, off, nd, _ = ens.slices[i]
        for c in range(3):
            assert relerr(grads[p][c][off:off + nd].cpu().numpy(), lam[c].cpu().numpy()) <= 1e-12
