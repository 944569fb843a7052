"""gf_integrate_noalloc: GGrid::integrate with its partial sums in the
library's fixed per-device scratch array instead of the mesh's reduction
buffer (for callers that cannot allocate).  Called through the C ABI on two
meshes in a row, the second much smaller than the first (so it uses fewer
scratch slots than the first call left filled), and checked bit for bit
against gf_integrate, which runs the same kernels with the same block count.
"""
import ctypes

import numpy as np
import pytest
import torch

import geoflow_amd as ga
from geoflow_amd import _capi, workloads
from geoflow_amd.device import Mesh

pytestmark = pytest.mark.gpu


def _noalloc(m, u):
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    _capi.check(_capi.lib().gf_integrate_noalloc(m.h, ctypes.c_void_p(u.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr())), m.ctx.h)
    torch.cuda.synchronize()
    return float(out.cpu().numpy()[0])


def test_integrate_noalloc_matches_integrate_on_two_calls(ctx):
    cases = [(ga.BoxGrid(3, 4, [16, 16, 16]), 11),   # 512,000 nodes: all 1024 slots
             (ga.BoxGrid(2, 4, [4, 4]), 12)]         # 400 nodes: 2 slots
    for bg, seed in cases:
        print("integrate_noalloc: ndof", bg.ndof, flush=True)
        m = Mesh(ctx, bg)
        u = torch.from_numpy(workloads.unique_normal(bg, seed=seed)).cuda()
        ref = m.integrate(u)
        got = _noalloc(m, u)
        assert got == ref, (bg.ndof, got, ref)
        # integral of 1 is the box volume (unit box)
        one = torch.ones(bg.ndof, dtype=torch.float64, device="cuda")
        assert abs(_noalloc(m, one) - 1.0) < 1e-13


def test_integrate_noalloc_rejects_null_arguments(ctx):
    bg = ga.BoxGrid(2, 3, [2, 2])
    m = Mesh(ctx, bg)
    rc = _capi.lib().gf_integrate_noalloc(m.h, None, None)
    assert rc == _capi.GF_ERR_ARG
