"""GGFX maps of a structured box built on the device (gf_box_ggfx_device,
and gf_mesh_create with box_device_maps): the CSR must equal the host
coordinate matching (gf_ggfx_build, the restatement of GGFX::init) entry for
entry, and a mesh built from device maps must give the same results bit for
bit as one built from host maps.  Boxes that mix periodic and non-periodic
directions are refused (GGridBox::periodize joins their groups differently).
"""
import numpy as np
import pytest
import torch

import geoflow_amd as ga
from geoflow_amd import _capi, workloads
from geoflow_amd.device import Mesh, box_ggfx_device

PER = [ga.GF_BDY_PERIODIC] * 6
DIR = [ga.GF_BDY_DIRICHLET] * 6


def test_device_maps_refuse_mixed_periodicity_on_host():
    with pytest.raises(ValueError):
        ga.BoxGrid(3, 3, [2, 2, 2], bdy=[ga.GF_BDY_PERIODIC] * 4 + DIR[:2], device_maps=True)
    with pytest.raises(ValueError):
        ga.BoxGrid(2, 3, [2, 2], device_maps=True)


@pytest.mark.gpu
@pytest.mark.parametrize("p,ne,bdy", [
    (1, [2, 3, 2], PER), (3, [3, 2, 4], PER), (7, [2, 2, 3], PER), (4, [1, 3, 2], PER),
    (3, [3, 2, 4], DIR), (7, [2, 3, 2], DIR), (2, [1, 1, 1], DIR), (5, [1, 2, 1], PER)])
def test_device_maps_equal_host_matching(ctx, p, ne, bdy):
    print("device GGFX", p, ne, bdy[0], flush=True)
    bg = ga.BoxGrid(3, p, ne, bdy=bdy)
    offs, mem = box_ggfx_device(ctx, p + 1, ne, bg.periodic)
    assert np.array_equal(offs, bg.ggfx_offs)
    assert np.array_equal(mem, bg.ggfx_members)


@pytest.mark.gpu
def test_device_maps_refuse_mixed_periodicity(ctx):
    with pytest.raises(_capi.GeoflowError) as ei:
        box_ggfx_device(ctx, 4, [2, 2, 2], [1, 1, 0])
    assert ei.value.code == _capi.GF_ERR_UNSUPPORTED


@pytest.mark.gpu
@pytest.mark.parametrize("bdy", [PER, DIR], ids=["periodic", "dirichlet"])
def test_mesh_from_device_maps_matches_host_maps(ctx, bdy):
    """Mass, inverse mass, multiplicity, DSS, a CG solve and (periodic) RK4
    steps equal bit for bit between meshes built from host and device maps."""
    print("mesh device maps", bdy[0], flush=True)
    hg = ga.BoxGrid(3, 7, [3, 2, 2], bdy=bdy)
    dg = ga.BoxGrid(3, 7, [3, 2, 2], bdy=bdy, device_maps=True)
    mh, md = Mesh(ctx, hg), Mesh(ctx, dg)
    assert mh.flags == md.flags
    for w in ("mass", "imass", "imult", "mask"):
        assert np.array_equal(mh.get(w), md.get(w)), w
    u = torch.from_numpy(workloads.unique_normal(hg, seed=3) + 0.1 * np.arange(hg.ndof) / hg.ndof)
    u = u.cuda()
    for op in ("sum", "smooth", "max", "min"):
        assert torch.equal(mh.dss(u.clone(), op), md.dss(u.clone(), op)), op
    b = mh.mass_apply(torch.from_numpy(workloads.unique_normal(hg, seed=4)).cuda())
    xb = torch.zeros_like(b) if bdy[0] == ga.GF_BDY_DIRICHLET else None
    for prec in (0, 1):
        xa, ia = mh.cg_solve(b, xb=xb, p=0.1, q=1.0, maxit=400, tol=1e-10, normtype=4, precond=1,
                             precision=prec)
        xd, idd = md.cg_solve(b, xb=xb, p=0.1, q=1.0, maxit=400, tol=1e-10, normtype=4, precond=1,
                              precision=prec)
        assert ia["iters"] == idd["iters"] and ia["status"] == idd["status"] == 0
        assert torch.equal(xa, xd)
    if bdy[0] == ga.GF_BDY_PERIODIC:
        comps = [0.5 * workloads.unique_normal(hg, seed=20 + c) for c in range(3)]
        ua = [torch.from_numpy(c).cuda() for c in comps]
        ud = [torch.from_numpy(c).cuda() for c in comps]
        mh.burgers_steps(ua, 0.02, 1e-3, 3)
        md.burgers_steps(ud, 0.02, 1e-3, 3)
        for a, d in zip(ua, ud):
            assert torch.equal(a, d)


@pytest.mark.parametrize("p,ne,bdy", [(3, [3, 2, 4], PER), (4, [1, 3, 2], PER), (2, [3, 2, 2], DIR),
                                      (1, [2, 2, 2], PER)])
def test_first_member_without_maps_equals_ggfx(p, ne, bdy):
    """The smallest group member formed from the box structure (used to draw
    per-unique-node random fields when the maps live on the device) equals
    the first member of the host GGFX CSR, so both grids get the same field."""
    from geoflow_amd.workloads import box_first_member
    hg = ga.BoxGrid(3, p, ne, bdy=bdy)
    dg = ga.BoxGrid(3, p, ne, bdy=bdy, device_maps=True)
    assert np.array_equal(box_first_member(dg), hg.ggfx_members[hg.ggfx_offs[:-1]])
    assert np.array_equal(workloads.unique_normal(dg, seed=5), workloads.unique_normal(hg, seed=5))
