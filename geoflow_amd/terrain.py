"""GGrid::add_terrain (src/cdg/grid/ggrid.ipp:1486-1548) for a 2-D box.

The reference builds the base grid of GE_DEFORMED elements, sets boundary
data xb from the base coordinates on the global boundary with the terrain
added to the bottom face (gterrainSpecBox, src/cdg/user/terrain/
gterrain_specbox_user.ipp:26-99: nodes within eps = minnodedist/8 of P0y, then
a GGFX Smooth of xb), solves for every coordinate the Laplace problem
L X = 0 with X = xb on the boundary (GCG with a GHelmholtz operator without
mass term, initial guess 0), installs the solutions as node coordinates and
recomputes the geometry.  Here the base grid is a flat TerrainBoxGrid
(Dirichlet on all four faces), the solves run on the device through
Mesh.cg_solve, and the geometry of the new coordinates comes from
gf_deformed_geometry2d.
"""
import numpy as np

from ._capi import GF_BDY_DIRICHLET
from .grid import DeformedGrid2D, TerrainBoxGrid


def terrain_boundary_data(base, terrain, P0y):
    """xb of gterrainSpecBox: base coordinates on boundary nodes (mask 0),
    zero inside, the terrain height added at the bottom face."""
    x, y = base.xyz
    bnd = base.mask == 0.0
    xb = [np.where(bnd, x, 0.0), np.where(bnd, y, 0.0)]
    eps = 0.125 * base.minnodedist
    bottom = bnd & (y <= P0y + eps) & (y >= P0y - eps)  # FUZZYEQ(P0.x2, y, eps)
    xb[1] = xb[1] + np.where(bottom, terrain(x), 0.0)
    return xb


def add_terrain(ctx, p, ne, terrain, P0=(0.0, 0.0), P1=(1.0, 1.0), reference_inverse=True,
                tol=1e-13, maxit=20000, normtype=0):
    """New GE_DEFORMED grid whose bottom follows terrain(x) (a vectorised
    height function); returns (grid, per-coordinate CG info)."""
    import torch
    from .device import Mesh
    base = TerrainBoxGrid(p, ne, amp=0.0, P0=P0, P1=P1, bdy=[GF_BDY_DIRICHLET] * 4,
                          reference_inverse=reference_inverse)
    m = Mesh(ctx, base)
    xb = terrain_boundary_data(base, terrain, P0[1])
    new, infos = [], []
    zero = torch.zeros(base.ndof, dtype=torch.float64, device="cuda")
    for j in range(2):
        xbj = m.dss(torch.from_numpy(xb[j]).cuda(), "smooth")
        x, info = m.cg_solve(zero, x=torch.zeros_like(zero), xb=xbj, p=1.0, q=0.0, maxit=maxit,
                             tol=tol, normtype=normtype)
        if info["status"] != 0:
            raise RuntimeError("add_terrain: CG did not converge for coordinate %d" % j)
        new.append(x.cpu().numpy())
        infos.append(info)
    return DeformedGrid2D(base, new, reference_inverse=reference_inverse), infos
