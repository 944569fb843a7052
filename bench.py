#!/usr/bin/env python
"""Benchmark of the GeoFLOW SEM operator path on one MI355X.

Workload: 3-D box of 32^3 hex elements at GLL order p=7 (16,777,216 local
DOF), the configuration the headline metric is quoted on:
  * Burgers RK4: nonlinear viscous Burgers, 3 velocity components, nu=0.02,
    dt=1e-5, fully periodic box; a "step" is one classical RK4 step of the
    whole state (GBurgers + GExRKStepper with GGFX Smooth).  `value` is
    DOF-steps/s = N_local * steps / seconds.
  * Helmholtz CG: (M + 0.1 L) u = b with homogeneous Dirichlet boundaries,
    b = DSS_sum(M f), f ~ N(0,1) per unique node; Jacobi PCG to assembled
    relative residual 1e-10, plain FP64 versus FP32 inner solves in FP64
    iterative refinement; plus 100 operator applies alone.
Inputs are synthetic and resident in HBM before timing starts.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "Burgers RK4 DOF-steps/s & CG solve s at rel-res 1e-10, mixed vs FP64"
HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel, profile="profiles/pmc_32cube_p7_kernels.md"):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE x2 + WRITE_SIZE, calibrated on a copy of known size), or None
    if the summary has no row for it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 4 and cells[1].strip("`").endswith(kernel):
            try:
                return (float(cells[2]) + float(cells[3])) * 1e6
            except ValueError:
                return None
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def burgers_ic(torch, xyz, seed=12345, nmodes=12):
    """Smooth periodic velocity: a few random Fourier modes per component."""
    rng = np.random.default_rng(seed)
    x, y, z = xyz
    comps = []
    for c in range(3):
        u = torch.zeros_like(x)
        for _ in range(nmodes):
            k = rng.integers(-3, 4, size=3)
            if not k.any():
                k[0] = 1
            a = rng.standard_normal() / (1.0 + (k * k).sum())
            ph = rng.uniform(0, 2 * np.pi)
            u += a * torch.cos(2 * np.pi * (int(k[0]) * x + int(k[1]) * y + int(k[2]) * z) + ph)
        comps.append(u.contiguous())
    return comps


def cpu_baseline(seconds=12.0):
    """Oracle (scalar C++ restatement of the reference) on a bounded sample of
    the same workload: 8^3 elements, p=7, periodic, 3 components; 1 thread
    (one MPI rank of the reference), then all of the host's cores (OpenMP
    over contiguous element ranges, as the reference's element decomposition
    over ranks would split the work)."""
    import oracle
    import geoflow_amd as ga
    og = oracle.Grid.box(3, 7, [8, 8, 8])
    bg = ga.BoxGrid(3, 7, [8, 8, 8])

    def run(threads, secs):
        oracle.set_threads(threads)
        us = [np.cos(2 * np.pi * (bg.xyz[0] + c * bg.xyz[1])) + 0.3 * np.sin(2 * np.pi * bg.xyz[2])
              for c in range(3)]
        og.burgers_step(us, 0.02, 1e-5)  # warm-up
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < secs:
            og.burgers_step(us, 0.02, 1e-5)
            n += 1
        return n, time.perf_counter() - t0

    n, dt = run(1, seconds)
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(ncores, 16))
    nm, dtm = run(ncores, max(2.0, seconds / 3))
    oracle.set_threads(1)
    return {"value": og.ndof * n / dt, "unit": "DOF-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle/gf_oracle.cpp (-O3 -march=native -ffp-contract=off), 1 thread: "
                      "%d RK4 Burgers steps of an 8^3-element p=7 periodic box (%d local DOF, "
                      "3 components) in %.1f s" % (n, og.ndof, dt),
            "all_cores": {"value": og.ndof * nm / dtm, "cores": ncores,
                          "sample": "same, OpenMP over %d threads: %d steps in %.1f s"
                                    % (ncores, nm, dtm)}}


def sphere_advection(ctx, level, p, nsteps, full=False, seed=12345, nu=0.0):
    """Cosine bell (radius R/3, seed-chosen centre) under Williamson test-1
    solid-body rotation tilted by pi/4 on the icosahedral sphere (GICOS
    bisection, GE_2DEMBEDDED quads), RK4 pure advection (nu = 0); one
    revolution takes nsteps steps.  short: times 100 steps; full: runs the
    revolution and reports L2/Linf errors against the exact (rotated) bell."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    from geoflow_amd import workloads as W
    t = time.perf_counter()
    g = ga.IcosGrid(p, level, radius=1.0)
    log("icosahedral grid level %d, p=%d: %d quads, %d local DOF, %.1f s"
        % (level, p, g.nelems, g.ndof, time.perf_counter() - t))
    m = Mesh(ctx, g)
    lat0, lon0 = W.bell_centre(seed)
    alpha, T = np.pi / 4, 1.0
    c0 = 2 * np.pi / T
    h0 = W.cosine_bell(*g.xyz, lat0, lon0)
    st = [torch.from_numpy(a).cuda() for a in [h0] + W.williamson1_velocity(*g.xyz, c0, alpha)]
    dt = T / nsteps
    m.burgers_step(st, nu, dt, mode=1)  # warm-up
    st[0].copy_(torch.from_numpy(h0))
    torch.cuda.synchronize()
    run = nsteps if full else min(100, nsteps)
    t0 = time.perf_counter()
    done = 0
    while done < run:
        k = min(1000, run - done)
        m.burgers_steps(st, nu, dt, k, t=done * dt, mode=1)
        done += k
        torch.cuda.synchronize()
        if full:
            log("sphere: %d / %d steps, %.1f s" % (done, run, time.perf_counter() - t0))
    el = time.perf_counter() - t0
    res = {"level": level, "p": p, "nu": nu, "quads": g.nelems, "ndof": g.ndof,
           "steps_per_revolution": nsteps,
           "dt": dt, "steps_timed": run, "ms_per_step": 1e3 * el / run,
           "dof_steps_per_s": g.ndof * run / el, "dtype": "f64",
           "bell_centre_deg": [float(np.rad2deg(lat0)), float(np.rad2deg(lon0))]}
    if full:
        h = st[0].cpu().numpy()
        ex = W.rotated_bell(*g.xyz, lat0, lon0, alpha, c0 * T)
        err = h - ex
        mass = m.get("mass")
        res["l2_error"] = float(np.sqrt(np.dot(mass, err * err) / np.dot(mass, ex * ex)))
        res["linf_error"] = float(np.abs(err).max() / np.abs(ex).max())
        log("sphere: L2 %.3e  Linf %.3e" % (res["l2_error"], res["linf_error"]))
    log("sphere advection: %.3f ms/step, %.3e DOF-steps/s" % (res["ms_per_step"], res["dof_steps_per_s"]))
    return res


def ensemble_run(ctx, nmembers, counts, nsteps=50, seed=12345, nu=0.02, dt=5e-5):
    """Batched ensemble (members drawn by seed from element counts `counts`
    per direction and orders {4, 6, 8}): nsteps implicit-viscous steps with
    CG sub-solves to rel-res 1e-10, FP32-inner/FP64-IR versus FP64, and the
    per-member error of the mixed run against the FP64 run."""
    import torch
    from geoflow_amd.ensemble import Ensemble, draw_members
    from geoflow_amd.workloads import fourier_field_3d
    members = draw_members(nmembers, seed=seed, counts=counts)
    t = time.perf_counter()
    ens = Ensemble(ctx, members)
    log("ensemble: %d members, %d local DOF, orders %s, set-up %.1f s"
        % (nmembers, ens.ndof, sorted(ens.mesh), time.perf_counter() - t))

    def fill(i, g):
        base = fourier_field_3d(*g.xyz, seed=seed, kmax=3)
        rng = np.random.default_rng(seed + 1 + i)
        return [b + 1e-3 * rng.standard_normal(g.ndof) for b in base]

    out = {"members": nmembers, "ndof": ens.ndof, "steps": nsteps, "nu": nu, "dt": dt,
           "orders": {str(p): len(ix) for p, ix in ens.groups.items()},
           "element_counts": sorted({m.ne for m in members})}
    finals = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        st = ens.new_state(fill)
        ens.imex_step(st, nu, dt, precision=prec)  # warm-up step
        st = ens.new_state(fill)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters = 0
        for _ in range(nsteps):
            its = ens.imex_step(st, nu, dt, precision=prec)
            iters += sum(sum(v) for v in its.values())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"seconds": el, "dof_steps_per_s": ens.ndof * nsteps / el, "cg_iters": iters}
        finals[name] = st
        log("ensemble %s: %.2f s, %.3e DOF-steps/s, %d CG iterations"
            % (name, el, ens.ndof * nsteps / el, iters))
    errs = []
    for i in range(nmembers):
        a, b = ens.member(finals["mixed"], i), ens.member(finals["fp64"], i)
        errs.append(max(float(np.abs(x - y).max() / np.abs(y).max()) for x, y in zip(a, b)))
    out["mixed_rel_err_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs))}
    out["speedup_mixed_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    return out


def kernel_set(ctx, ne, p, reps):
    """Average time (ms, HIP events) of each kernel of the path on the
    periodic (Burgers) and Dirichlet (CG) boxes, after a plain device copy of
    known size that profiler counter passes use for calibration."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    a = torch.ones(ne ** 3 * (p + 1) ** 3, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    del a, b
    res = {}
    for kind in ("periodic", "dirichlet"):
        bdy = [ga.GF_BDY_DIRICHLET] * 6 if kind == "dirichlet" else None
        m = Mesh(ctx, ga.BoxGrid(3, p, [ne] * 3, bdy=bdy))
        names = ("burgers_stage_box", "rk_out_box", "helmholtz") if kind == "periodic" else \
                ("cg_apply", "cg_apply_f32", "cg_update", "cg_update_f32")
        for k in names:
            res[k] = m.time_kernel(k, reps)
            log("%s %s %.1f us" % (kind, k, res[k] * 1e3))
        del m
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ne", type=int, default=32)
    ap.add_argument("--p", type=int, default=7)
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--sphere", choices=["off", "short", "full"], default="short",
                    help="icosahedral sphere advection: short = timed steps only, "
                         "full = one revolution with L2/Linf errors")
    ap.add_argument("--sphere-level", type=int, default=6)
    ap.add_argument("--sphere-steps", type=int, default=10000, help="RK4 steps per revolution")
    ap.add_argument("--sphere-nu", type=float, default=0.0,
                    help="viscosity of the advected field (0: pure advection)")
    ap.add_argument("--ensemble", type=int, default=8,
                    help="members of the batched ensemble (0: off)")
    ap.add_argument("--ensemble-counts", default="8,12",
                    help="element counts per direction to draw from (full workload: 8,12,16,24)")
    ap.add_argument("--kernels", type=int, default=0, metavar="REPS",
                    help="only time the path's kernels REPS times each on the 32^3 p=7 boxes "
                         "(for profiler counter passes) and print their average times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(see DESIGN.md for how the path would shard)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Context, Mesh
    from geoflow_amd import workloads

    ne, p = args.ne, args.p
    ctx = Context(0)
    if args.kernels:
        print(json.dumps(kernel_set(ctx, ne, p, args.kernels)), flush=True)
        return
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic"}

    # ---------------- Burgers RK4 ------------------------------------------
    t = time.perf_counter()
    bg = ga.BoxGrid(3, p, [ne] * 3)
    log("periodic box grid + GGFX maps: %.1f s, ndof=%d" % (time.perf_counter() - t, bg.ndof))
    m = Mesh(ctx, bg)
    ndof = m.ndof
    xyz = [torch.from_numpy(a).cuda() for a in bg.xyz]
    us = burgers_ic(torch, xyz)
    del xyz
    nu, dt = 0.02, 1e-5
    for _ in range(args.warmup):
        m.burgers_step(us, nu, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.burgers_steps(us, nu, dt, args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ok = all(bool(torch.isfinite(u).all()) for u in us)
    dofsteps = ndof * args.steps / el
    log("Burgers RK4: %d steps in %.3f s -> %.3e DOF-steps/s (finite=%s)" % (args.steps, el, dofsteps, ok))
    # the stage kernel the step runs: the structured-box form (previous K
    # assembled by gather inside the stage) when the mesh allows it
    boxed = bool(m.flags & ga._capi.GF_MESH_BOX_GATHER)
    stage_ms = m.time_kernel("burgers_stage_box" if boxed else "burgers_stage", reps=20)
    stage_name = ("k_burgers_stage3b<8> (structured-box RK stage, K gathered in-kernel; "
                  "middle-stage form)" if boxed else
                  "k_burgers_stage3p<8> (fused RK stage; middle-stage form)")
    stage_bytes = 5 * 8 * 3 * ndof  # middle-stage form: read uin,kprev,acc; write acc,k (per comp)
    stage_gbs = stage_bytes / (stage_ms * 1e-3) / 1e9
    step_bytes = 16 * 8 * 3 * ndof
    step_gbs = step_bytes / (el / args.steps) / 1e9
    out.update({"value": dofsteps, "unit": "DOF-steps/s", "ms_per_step": 1e3 * el / args.steps})
    traffic = measured_traffic("k_burgers_stage3b<%d, true, 2, false>" % (p + 1)) if boxed else None
    out["roofline"] = {"kernel": stage_name,
                       "bound": "hbm", "achieved": stage_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": stage_gbs / HBM_PEAK_GBS, "traffic": traffic,
                       "traffic_source": "profiles/pmc_32cube_p7_kernels.md" if traffic else None,
                       "avg_launch_ms": stage_ms, "alg_bytes_per_launch": stage_bytes,
                       "step_alg_bytes": step_bytes, "step_achieved_gbs": step_gbs,
                       "step_frac": step_gbs / HBM_PEAK_GBS}
    del us, m
    torch.cuda.empty_cache()

    # ---------------- Helmholtz CG -------------------------------------------
    if not args.no_cg:
        t = time.perf_counter()
        cg_grid = ga.BoxGrid(3, p, [ne] * 3, bdy=[ga.GF_BDY_DIRICHLET] * 6)
        log("Dirichlet box grid + GGFX maps: %.1f s" % (time.perf_counter() - t))
        m = Mesh(ctx, cg_grid)
        f = torch.from_numpy(workloads.unique_normal(cg_grid, seed=67890)).cuda()
        b = m.mass_apply(f)  # local RHS; GCG assembles b - A x (so the system's RHS is DSS_sum(M f))
        q = 1.0  # (M + 0.1 L): a one-entry q of 1, as GHelmholtz::set_mass_scalar allows
        mask = torch.from_numpy(cg_grid.mask).cuda()
        zero = torch.zeros_like(b)
        bA = m.dss(b.clone(), "sum") * mask
        bnorm = float(torch.sqrt((bA ** 2 * m.imult()).sum()))

        def true_relres(x):
            r = b - m.helmholtz(x, 0.1, q)
            m.dss(r, "sum")
            r *= mask
            return float(torch.sqrt((r * r * m.imult()).sum())) / bnorm

        res = {}
        for name, prec in (("fp64", 0), ("mixed", 1)):
            x = torch.zeros_like(b)
            m.cg_solve(b, x=x, xb=zero, p=0.1, q=q, maxit=30, tol=1e-10, normtype=4, precond=1,
                       precision=prec)  # warm-up (allocations, code objects)
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, info = m.cg_solve(b, x=x, xb=zero, p=0.1, q=q, maxit=20000, tol=1e-10, normtype=4,
                                 precond=1, precision=prec)
            torch.cuda.synchronize()
            el_cg = time.perf_counter() - t0
            rr = true_relres(x)
            res[name] = {"solve_s": el_cg, "iters": info["iters"], "outer": info["outer"],
                         "status": info["status"], "true_relres": rr}
            log("CG %s: %.3f s, %d iters (outer %d), rel-res %.2e, status %d"
                % (name, el_cg, info["iters"], info["outer"], rr, info["status"]))
            if name == "fp64":
                xref = x.clone()
            else:
                res[name]["x_rel_err_vs_fp64"] = float((x - xref).abs().max() / xref.abs().max())
        u = torch.randn(ndof, dtype=torch.float64, device="cuda")
        o = torch.empty_like(u)
        m.helmholtz(u, 0.1, q, out=o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(100):
            m.helmholtz(u, 0.1, q, out=o)
        torch.cuda.synchronize()
        a100 = time.perf_counter() - t0
        ap_ms = m.time_kernel("cg_apply", reps=20)
        ap32_ms = m.time_kernel("cg_apply_f32", reps=20)
        res["apply100_s"] = a100
        res["apply_dof_per_s"] = 100 * ndof / a100
        res["speedup_mixed_vs_fp64"] = res["fp64"]["solve_s"] / res["mixed"]["solve_s"]
        res["cg_apply_fp64"] = {"avg_launch_ms": ap_ms, "alg_bytes_per_launch": 5 * 8 * ndof,
                                "achieved_gbs": 5 * 8 * ndof / (ap_ms * 1e-3) / 1e9}
        res["cg_apply_fp64"]["frac"] = res["cg_apply_fp64"]["achieved_gbs"] / HBM_PEAK_GBS
        res["cg_apply_fp32"] = {"avg_launch_ms": ap32_ms, "alg_bytes_per_launch": 5 * 4 * ndof,
                                "achieved_gbs": 5 * 4 * ndof / (ap32_ms * 1e-3) / 1e9}
        res["cg_apply_fp32"]["frac"] = res["cg_apply_fp32"]["achieved_gbs"] / HBM_PEAK_GBS
        log("100 applies %.3f s; cg_apply fp64 %.3f ms, fp32 %.3f ms; speedup %.2fx"
            % (a100, ap_ms, ap32_ms, res["speedup_mixed_vs_fp64"]))
        out["cg"] = res
        del m

    if args.ensemble > 0:
        out["ensemble"] = ensemble_run(ctx, args.ensemble,
                                       tuple(int(c) for c in args.ensemble_counts.split(",")))
    if args.sphere != "off":
        out["sphere"] = sphere_advection(ctx, args.sphere_level, 5, args.sphere_steps,
                                         full=args.sphere == "full", nu=args.sphere_nu)

    out["config"] = {"workload": "3-D box %d^3 hex elements, p=%d (%d local DOF): Burgers RK4 "
                                 "(3 comps, nu=0.02, dt=1e-5, periodic) and Helmholtz (M+0.1L) "
                                 "Jacobi-PCG to rel-res 1e-10 (Dirichlet), FP64 vs FP32-inner/FP64-IR"
                                 % (ne, p, ndof),
                     "elements": [ne, ne, ne], "p": p, "ndof": ndof, "global_batch": ndof,
                     "parallelism": "single GPU"}
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        log("CPU baseline %.3e DOF-steps/s" % out["cpu_baseline"]["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
