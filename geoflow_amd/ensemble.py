"""Batched ensembles of Burgers members with uneven element counts and orders.

Members are periodic boxes [0,1]^3 of their own element count and GLL order.
Members of one order are concatenated into a single device mesh whose GGFX
groups are block-diagonal (a group never crosses members), so every element
kernel, DSS pass and CG kernel advances all members of that order in one
launch; an ensemble with orders {4, 6, 8} needs three launches per kernel,
whatever the number of members.

The step is implicit in the viscous term (the reference has no implicit
stepper; this implicit form is an addition for the ensemble workload):
    M u*   = M u^n - dt M (u^n . grad) u^n           (GAdvect, explicit)
    (M + nu dt L) u^{n+1} = M u*                     (Helmholtz, Jacobi PCG)
per velocity component, the solve in FP64 or with FP32 inner iterations and
FP64 reliable updates.  The CG of a group runs all its members in the same
kernel launches, each member with its own CG scalars, norm of b and stopping
test (gf_mesh_set_segments), so every member reaches the relative residual
on its own system.

The ensemble gradient of a misfit J = 1/2 sum ||u(T) - u_target||_M^2 is
taken through the explicit RK4 stepper's adjoint (gf_burgers_adjoint) on the
same concatenated meshes: members do not interact, so each member's slice of
the gradient is that member's gradient.
"""
import numpy as np

from .grid import BoxGrid, gll_basis


class MemberSpec:
    def __init__(self, p, ne):
        self.p, self.ne = int(p), int(ne)
        self.ndof = self.ne ** 3 * (self.p + 1) ** 3


def draw_members(nmembers=48, seed=12345, counts=(8, 12, 16, 24), orders=(4, 6, 8)):
    """Element counts per direction and orders drawn by seed."""
    rng = np.random.default_rng(seed)
    return [MemberSpec(int(rng.choice(orders)), int(rng.choice(counts))) for _ in range(nmembers)]


class _UnionGrid:
    """Host arrays of several periodic boxes of one order, concatenated.
    With device_maps the boxes' GGFX groups are built on the device at mesh
    creation (box_list describes the boxes) and no host maps are formed."""

    def __init__(self, grids, device_maps=False):
        g0 = grids[0]
        self.dim = 3
        self.gtype = g0.gtype
        self.p = g0.p
        self.N = g0.N
        self.basis = g0.basis
        self.nelems = sum(g.nelems for g in grids)
        self.ndof = sum(g.ndof for g in grids)
        # coordinates only where the host matching needs them (regular boxes'
        # operators do not read them)
        self.xyz = None if device_maps else [np.concatenate([g.xyz[d] for g in grids])
                                             for d in range(3)]
        self.edx = np.concatenate([g.edx for g in grids])
        self.ejac = np.concatenate([g.ejac for g in grids])
        self.mask = None
        self.ne = None  # no single box structure: CSR DSS path
        self.device_maps = bool(device_maps)
        if self.device_maps:
            self.box_list = [v for g in grids for v in list(g.ne) + list(g.periodic)]
            self.ggfx_offs = self.ggfx_members = None
            return
        offs, mem, base = [np.zeros(1, dtype=np.int64)], [], 0
        for g in grids:
            offs.append(g.ggfx_offs[1:] + (offs[-1][-1] if len(offs) > 1 else 0))
            mem.append(g.ggfx_members + base)
            base += g.ndof
        self.ggfx_offs = np.concatenate(offs)
        self.ggfx_members = np.concatenate(mem)


class Ensemble:
    """Members grouped by order; one device mesh per group."""

    def __init__(self, ctx, members, device_maps=False):
        """device_maps: build the members' GGFX groups on the device (no host
        coordinate matching; the fast set-up for large ensembles)."""
        from .device import Mesh
        self.ctx = ctx
        self.members = list(members)
        self.groups = {}
        for idx, ms in enumerate(self.members):
            self.groups.setdefault(ms.p, []).append(idx)
        self.mesh, self.slices, self.grids = {}, {}, {}
        for p, idxs in sorted(self.groups.items()):
            grids = [BoxGrid(3, p, [self.members[i].ne] * 3, device_maps=device_maps)
                     for i in idxs]
            ug = _UnionGrid(grids, device_maps=device_maps)
            self.mesh[p] = Mesh(ctx, ug, structured=False)
            # each member its own CG scalars and stopping test
            self.mesh[p].set_segments(np.cumsum([0] + [g.nelems for g in grids]))
            self.grids[p] = ug
            off = 0
            for i, g in zip(idxs, grids):
                self.slices[i] = (p, off, g.ndof, g)
                off += g.ndof

    @property
    def ndof(self):
        return sum(m.ndof for m in self.mesh.values())

    def new_state(self, fill):
        """State per group: 3 device fields; fill(i, grid) gives member i's
        3 fields (host arrays or device tensors)."""
        import torch
        st = {}
        for p, idxs in self.groups.items():
            comps = [torch.zeros(self.mesh[p].ndof, dtype=torch.float64, device="cuda")
                     for _ in range(3)]
            for i in idxs:
                _, off, n, g = self.slices[i]
                vals = fill(i, g)
                for c in range(3):
                    comps[c][off:off + n] = torch.as_tensor(vals[c], dtype=torch.float64,
                                                            device="cuda")
            st[p] = comps
        return st

    def member(self, st, i):
        """Host copies of member i's 3 components."""
        p, off, n, _ = self.slices[i]
        return [st[p][c][off:off + n].cpu().numpy() for c in range(3)]

    def imex_step(self, st, nu, dt, precision=0, tol=1e-10, maxit=2000, inner_tol=1e-5):
        """One implicit-viscous step of every member; returns the CG
        iteration counts, per group and component.  The sub-solves take ~5
        iterations, so the FP32-inner solver refines only when its residual
        has fallen by inner_tol (1e-5: one or two FP64 reliable updates per
        solve; the solver's default 0.01 would need five)."""
        import torch
        its = {}
        self.last_member_iters = {}
        for p, m in self.mesh.items():
            us = st[p]
            new = []
            for c in range(3):
                b = m.mass_apply(us[c])
                b -= dt * m.advect(us[c], us)
                x = us[c].clone()
                x, info = m.cg_solve(b, x=x, p=nu * dt, q=1.0, maxit=maxit, tol=tol, normtype=4,
                                     precond=1, precision=precision, inner_tol=inner_tol)
                if info["status"] != 0:
                    raise RuntimeError("ensemble CG did not converge (order %d, comp %d)" % (p, c))
                its.setdefault(p, []).append(info["iters"])
                self.last_member_iters.setdefault(p, []).append(m.member_results()[0])
                new.append(x)
            for c in range(3):
                us[c].copy_(new[c])
        torch.cuda.synchronize()
        return its

    def rk4_gradient(self, st0, targets, nu, dt, nsteps, ckpt=10):
        """Gradient of J = 1/2 sum_members ||u(T) - target||_M^2 with respect
        to u(0) for nsteps explicit RK4 steps; targets like st0."""
        grads, uT = {}, {}
        for p, m in self.mesh.items():
            fw = [a.clone() for a in st0[p]]
            m.burgers_steps(fw, nu, dt, nsteps)
            mass = m.new()
            mass.copy_(__import__("torch").from_numpy(m.get("mass")).cuda())
            lam = [mass * (a - t) for a, t in zip(fw, targets[p])]
            m.burgers_adjoint([a.clone() for a in st0[p]], lam, nu, dt, nsteps, ckpt=ckpt)
            grads[p], uT[p] = lam, fw
        return grads, uT
