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
import oracle

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


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_batched_members_converge_each_and_match_oracle(ctx, precision):
    """Members of orders 4, 6 and 8 in one batch, each member with its own
    CG scalars and stopping test: every member's assembled residual reaches
    1e-10 of its own right-hand side, and every member's implicit step equals
    the oracle's step (the oracle's GAdvect for the explicit part, its GCG
    with the Jacobi diagonal for (M + nu dt L) x = b, solved to 1e-13) to
    1e-8, the accuracy a 1e-10 residual gives on this well-conditioned
    system."""
    from geoflow_amd.ensemble import Ensemble
    import torch
    print("ensemble per-member CG", precision, flush=True)
    members = [MemberSpec(4, 2), MemberSpec(6, 2), MemberSpec(8, 2), MemberSpec(4, 4),
               MemberSpec(6, 2)]
    ens = Ensemble(ctx, members)
    st = ens.new_state(_ic(21))
    nu, dt = 0.05, 2e-3
    ens.imex_step(st, nu, dt, precision=precision, tol=1e-10)
    for p, m in ens.mesh.items():
        its, rn, r0 = m.member_results()
        assert (rn <= 1e-10 * r0).all(), (p, rn / r0)
    for i, ms in enumerate(members):
        og = oracle.Grid.box(3, ms.p, [ms.ne] * 3)
        g = ga.BoxGrid(3, ms.p, [ms.ne] * 3)
        us = _ic(21)(i, g)
        got = ens.member(st, i)
        for c in range(3):
            b = og.mass_apply(us[c]) - dt * og.advect(us[c], us)
            x, ret, it, _ = og.cg_solve(b, x0=us[c], p=nu * dt, q=np.ones(og.ndof), maxit=2000,
                                        tol=1e-13, normtype=4, precond=1)
            assert ret == 0
            assert relerr(got[c], x) <= 1e-8, (i, c, relerr(got[c], x))


@pytest.mark.gpu
def test_member_offsets_must_not_split_groups(ctx):
    """Members may start at any element (at p=4 the CG apply works two
    elements per workgroup, and a workgroup may then hold two members), but
    no GGFX group may join two members."""
    from geoflow_amd import _capi
    from geoflow_amd.device import Mesh
    grids = [ga.BoxGrid(3, 4, [3, 3, 3]), ga.BoxGrid(3, 4, [2, 2, 2])]
    m = Mesh(ctx, _UnionGrid(grids), structured=False)
    m.set_segments([0, 27, 35])
    g2 = [ga.BoxGrid(3, 4, [2, 2, 2]), ga.BoxGrid(3, 4, [2, 2, 2])]
    m2 = Mesh(ctx, _UnionGrid(g2), structured=False)
    with pytest.raises(_capi.GeoflowError):
        m2.set_segments([0, 4, 16])
    m2.set_segments([0, 8, 16])


@pytest.mark.gpu
def test_device_built_member_maps_equal_host_maps(ctx):
    """An ensemble whose members' GGFX groups are built on the device (the
    union of boxes described to gf_mesh_create) steps bit for bit like the
    one built from the host's coordinate matching."""
    from geoflow_amd.ensemble import Ensemble
    import torch
    print("ensemble device maps", flush=True)
    members = [MemberSpec(4, 2), MemberSpec(4, 4), MemberSpec(6, 2), MemberSpec(8, 2)]
    eh = Ensemble(ctx, members)
    ed = Ensemble(ctx, members, device_maps=True)
    for p in eh.mesh:
        for w in ("mass", "imass", "imult"):
            assert np.array_equal(eh.mesh[p].get(w), ed.mesh[p].get(w)), (p, w)
    sh, sd = eh.new_state(_ic(31)), ed.new_state(_ic(31))
    eh.imex_step(sh, 0.05, 2e-3, precision=1)
    ed.imex_step(sd, 0.05, 2e-3, precision=1)
    for p in sh:
        for a, b in zip(sh[p], sd[p]):
            assert torch.equal(a, b)
