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
import subprocess
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


def host_info():
    """Host CPU as lscpu and nproc report it."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                           timeout=20).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        info["nproc"] = os.cpu_count() or 1
    return info


def median_time(fn, warm=3, runs=5):
    """Wall clock of fn(): median of `runs` runs after `warm` warm-up runs."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline():
    """Oracle (C++ restatement of the reference, -O3 -march=native
    -ffp-contract=off) on bounded samples of configs 1-2 and 4 of the
    workload set, on 1 thread (one MPI rank of the reference) and on the
    host's threads (OpenMP over contiguous index ranges of every
    element-local and pointwise pass, as the reference's element
    decomposition over ranks splits the work; dot products stay serial).
    Each figure is the median of 5 runs after 3 warm-up runs; set-up (grid,
    GGFX maps, basis) is excluded."""
    import oracle
    import geoflow_amd as ga
    from geoflow_amd import workloads
    info = host_info()
    nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["nproc"]
    # RK4 Burgers: 8^3 elements, p=7, periodic, 3 components; a run is 10 steps
    og = oracle.Grid.box(3, 7, [8, 8, 8])
    bg = ga.BoxGrid(3, 7, [8, 8, 8])
    us = [np.cos(2 * np.pi * (bg.xyz[0] + c * bg.xyz[1])) + 0.3 * np.sin(2 * np.pi * bg.xyz[2])
          for c in range(3)]
    # FP64 Jacobi-PCG: 4^3 elements, p=7, Dirichlet, (M + 0.1 L) u = DSS(M f)
    dirich = [ga.GF_BDY_DIRICHLET] * 6
    ocg = oracle.Grid.box(3, 7, [4, 4, 4], bdy=dirich)
    bcg = ga.BoxGrid(3, 7, [4, 4, 4], bdy=dirich)
    bvec = ocg.mass_apply(workloads.unique_normal(bcg, seed=67890))
    ones, zero = np.ones(ocg.ndof), np.zeros(ocg.ndof)
    cg_it = [0]
    # forward + adjoint: 4^3 elements, p=6, periodic, nu=0.02, dt=5e-5, a run
    # is 10 RK4 steps forward (states kept) and their 10 adjoint steps
    oad = oracle.Grid.box(3, 6, [4, 4, 4])
    bad = ga.BoxGrid(3, 6, [4, 4, 4])
    u0 = workloads.fourier_field_3d(*bad.xyz, seed=12345, kmax=6)

    def burgers():
        for _ in range(10):
            og.burgers_step(us, 0.02, 1e-5)

    def cg():
        _, _, it, _ = ocg.cg_solve(bvec, xb=zero, p=0.1, q=ones, maxit=5000, tol=1e-10,
                                   normtype=4, precond=1)
        cg_it[0] = it

    def adjoint():
        u = [a.copy() for a in u0]
        states = []
        for _ in range(10):
            states.append([a.copy() for a in u])
            oad.burgers_step(u, 0.02, 5e-5)
        lam = [a.copy() for a in u]
        for st in reversed(states):
            oad.burgers_step_adjoint(st, lam, 0.02, 5e-5)

    res = {}
    for thr in (1, nthr):
        oracle.set_threads(thr)
        tb = median_time(burgers)
        tc = median_time(cg)
        ta = median_time(adjoint)
        res[thr] = {"burgers": og.ndof * 10 / tb, "cg_s": tc, "cg_iters": cg_it[0],
                    "cg_dof_iters": ocg.ndof * cg_it[0] / tc, "adjoint": bad.ndof * 10 / ta}
        log("CPU baseline, %d thread(s): Burgers %.3e DOF-steps/s, CG %.3f s (%d it), "
            "forward+adjoint %.3e DOF-steps/s" % (thr, res[thr]["burgers"], tc, cg_it[0],
                                                  res[thr]["adjoint"]))
    oracle.set_threads(1)
    r1, rn = res[1], res[nthr]
    return {"value": r1["burgers"], "unit": "DOF-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle/gf_oracle.cpp (-O3 -march=native -ffp-contract=off), 1 thread: "
                      "RK4 Burgers steps of an 8^3-element p=7 periodic box (%d local DOF, "
                      "3 components), median of 5 runs of 10 steps after 3 warm-up runs"
                      % og.ndof,
            "host": info,
            "all_cores": {"value": rn["burgers"], "cores": nthr,
                          "sample": "same, OpenMP over %d threads" % nthr},
            "cg": {"sample": "FP64 Jacobi-PCG (M + 0.1 L) u = DSS(M f), 4^3 elements, p=7, "
                             "Dirichlet, to assembled rel-res 1e-10 (%d local DOF)" % ocg.ndof,
                   "iters": r1["cg_iters"], "solve_s": r1["cg_s"],
                   "dof_iters_per_s": r1["cg_dof_iters"],
                   "all_cores": {"cores": nthr, "solve_s": rn["cg_s"],
                                 "dof_iters_per_s": rn["cg_dof_iters"]}},
            "adjoint": {"sample": "10 RK4 steps forward (nu=0.02, dt=5e-5) and their adjoint, "
                                  "4^3 elements, p=6, periodic (%d local DOF)" % bad.ndof,
                        "dof_steps_per_s": r1["adjoint"],
                        "all_cores": {"cores": nthr, "dof_steps_per_s": rn["adjoint"]}}}


def sphere_advection(ctx, level, p, nsteps, full=False, seed=12345, nu=0.0):
    """Cosine bell (radius R/3, seed-chosen centre) under Williamson test-1
    solid-body rotation tilted by pi/4 on the icosahedral sphere (GICOS
    bisection, GE_2DEMBEDDED quads), RK4 advection with viscosity nu; one
    revolution takes nsteps steps.  FP64 and mixed precision (FP32 field and
    stage fields, FP64 metric terms).  short: times 100 steps of each; full:
    runs the revolution in each precision and reports L2/Linf errors against
    the exact (rotated, inviscid) bell and the mixed run's difference from
    the FP64 run."""
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
    vel = [torch.from_numpy(a).cuda() for a in W.williamson1_velocity(*g.xyz, c0, alpha)]
    dt = T / nsteps
    run = nsteps if full else min(100, nsteps)
    res = {"level": level, "p": p, "nu": nu, "quads": g.nelems, "ndof": g.ndof,
           "steps_per_revolution": nsteps, "dt": dt, "steps_timed": run,
           "bell_centre_deg": [float(np.rad2deg(lat0)), float(np.rad2deg(lon0))]}
    mass = m.get("mass")
    ex = W.rotated_bell(*g.xyz, lat0, lon0, alpha, c0 * T) if full else None
    finals = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        st = [torch.from_numpy(h0).cuda()] + vel
        m.burgers_steps(st, nu, dt, 1, mode=1, precision=prec)  # warm-up
        st[0].copy_(torch.from_numpy(h0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < run:
            k = min(1000, run - done)
            m.burgers_steps(st, nu, dt, k, t=done * dt, mode=1, precision=prec)
            done += k
            torch.cuda.synchronize()
            if full:
                log("sphere %s: %d / %d steps, %.1f s" % (name, done, run, time.perf_counter() - t0))
        el = time.perf_counter() - t0
        r = {"ms_per_step": 1e3 * el / run, "dof_steps_per_s": g.ndof * run / el,
             "dtype": "f64" if prec == 0 else "f32 field / f64 metric"}
        h = st[0].cpu().numpy()
        finals[name] = h
        if full:
            err = h - ex
            r["l2_error"] = float(np.sqrt(np.dot(mass, err * err) / np.dot(mass, ex * ex)))
            r["linf_error"] = float(np.abs(err).max() / np.abs(ex).max())
            log("sphere %s: L2 %.3e  Linf %.3e" % (name, r["l2_error"], r["linf_error"]))
        log("sphere advection %s: %.3f ms/step, %.3e DOF-steps/s"
            % (name, r["ms_per_step"], r["dof_steps_per_s"]))
        res[name] = r
    d = finals["mixed"] - finals["fp64"]
    res["mixed_vs_fp64"] = {"linf_rel": float(np.abs(d).max() / np.abs(finals["fp64"]).max()),
                            "l2_rel": float(np.sqrt(np.dot(mass, d * d) /
                                                    np.dot(mass, finals["fp64"] ** 2)))}
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
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
    ens = Ensemble(ctx, members, device_maps=True)
    log("ensemble: %d members, %d local DOF, orders %s, set-up %.1f s"
        % (nmembers, ens.ndof, sorted(ens.mesh), time.perf_counter() - t))

    cache = {}

    def fill(i, g):
        # the config-4 Fourier field (formed on the device) plus 1e-3 N(0,1)
        # noise per member, drawn once and reused by both runs
        if i not in cache:
            xyz = [torch.from_numpy(a).cuda() for a in g.xyz]
            base = fourier_field_3d(*xyz, seed=seed, kmax=3, xp=torch)
            gen = torch.Generator(device="cuda").manual_seed(seed + 1 + i)
            cache[i] = [b + 1e-3 * torch.randn(g.ndof, dtype=torch.float64, device="cuda",
                                               generator=gen) for b in base]
        return cache[i]

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
        mi = [int(a.max()) for v in ens.last_member_iters.values() for a in v]
        out[name] = {"seconds": el, "dof_steps_per_s": ens.ndof * nsteps / el, "cg_iters": iters,
                     "last_step_member_iters_max": max(mi) if mi else 0}
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


def copy_peak(torch, nbytes=2 * 134217728, reps=20):
    """Measured HBM bandwidth of a plain device copy (torch copy_, read +
    write counted), timed with HIP events on the current stream."""
    a = torch.ones(nbytes // 16, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * (nbytes // 2) / (ms * 1e-3) / 1e9


def adjoint_run(ctx, ne=16, p=6, nsteps=200, ckpt=20, nu=0.02, dt=5e-5):
    """Config 4 on the GPU: forward RK4 Burgers and the checkpointed adjoint
    gradient of 0.5 ||u(T) - u_target||_M^2 with respect to u(0) on a 16^3
    periodic box at p=6, 200 steps of dt=5e-5, checkpoints every 20 steps."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    from geoflow_amd.workloads import fourier_field_3d
    g = ga.BoxGrid(3, p, [ne] * 3)
    m = Mesh(ctx, g)
    u0 = [torch.from_numpy(a).cuda() for a in fourier_field_3d(*g.xyz, seed=12345, kmax=6)]
    ut = [torch.from_numpy(a).cuda() for a in fourier_field_3d(*g.xyz, seed=2, kmax=6)]
    mass = torch.from_numpy(m.get("mass")).cuda()

    def fwd():
        u = [a.clone() for a in u0]
        m.burgers_steps(u, nu, dt, nsteps)
        return u

    def grad():
        # dJ/du(T) = M (u(T) - u_target) needs u(T): one forward sweep, then
        # gf_burgers_adjoint (its checkpointed forward sweep and the reverse
        # sweep that recomputes each segment)
        u = fwd()
        lam = [mass * (a - b) for a, b in zip(u, ut)]
        m.burgers_adjoint(u0, lam, nu, dt, nsteps, ckpt=ckpt)
        return lam

    grad()  # warm-up (workspace, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd()
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    lam = grad()
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(a).all()) for a in lam)
    log("adjoint 16^3 p=6: forward %.3f s, forward + gradient %.3f s" % (tf, tg))
    return {"elements": [ne] * 3, "p": p, "ndof": g.ndof, "steps": nsteps, "ckpt": ckpt,
            "nu": nu, "dt": dt, "forward_s": tf, "forward_and_gradient_s": tg,
            "dof_steps_per_s_forward": g.ndof * nsteps / tf, "finite": finite}


def node_metric_cg(ctx, runs=3):
    """FP64 versus mixed-precision Jacobi-PCG of (M + 0.1 L) u = DSS(M f) to
    assembled rel-res 1e-10 on meshes with per-node metrics: the icosahedral
    sphere (level 5, p=5) and the 3-D terrain-following box (16x16x8, p=7).
    Median of `runs` solves after one warm-up solve; the true FP64 assembled
    relative residual of each returned solution."""
    import torch
    import geoflow_amd as ga
    from geoflow_amd.device import Mesh
    from geoflow_amd import workloads
    out = {}
    for name, mk in (("sphere_l5_p5", lambda: ga.IcosGrid(5, 5, radius=1.0)),
                     ("terrain3d_16x16x8_p7", lambda: ga.TerrainBoxGrid3D(7, [16, 16, 8],
                                                                          amp=0.1))):
        g = mk()
        m = Mesh(ctx, g)
        b = m.mass_apply(torch.from_numpy(workloads.unique_normal(g, seed=67890)).cuda())
        has_mask = getattr(g, "mask", None) is not None and float(np.min(g.mask)) == 0.0
        mask = torch.from_numpy(g.mask).cuda() if has_mask else None
        xb = torch.zeros_like(b) if has_mask else None
        imult = m.imult()
        bA = m.dss(b.clone(), "sum")
        if mask is not None:
            bA *= mask
        bn = float(torch.sqrt((bA * bA * imult).sum()))
        res = {"ndof": g.ndof}
        for pname, prec in (("fp64", 0), ("mixed", 1)):
            def solve():
                x, info = m.cg_solve(b, xb=xb, p=0.1, q=1.0, maxit=20000, tol=1e-10, normtype=4,
                                     precond=1, precision=prec)
                torch.cuda.synchronize()
                return x, info
            solve()
            ts = []
            for _ in range(runs):
                t0 = time.perf_counter()
                x, info = solve()
                ts.append(time.perf_counter() - t0)
            r = b - m.helmholtz(x, 0.1, 1.0)
            m.dss(r, "sum")
            if mask is not None:
                r *= mask
            rr = float(torch.sqrt((r * r * imult).sum())) / bn
            res[pname] = {"solve_s": float(np.median(ts)), "iters": info["iters"],
                          "outer": info["outer"], "status": info["status"], "true_relres": rr}
        res["speedup_mixed_vs_fp64"] = res["fp64"]["solve_s"] / res["mixed"]["solve_s"]
        log("node-metric CG %s: fp64 %.4f s (%d it), mixed %.4f s (%d it), %.2fx"
            % (name, res["fp64"]["solve_s"], res["fp64"]["iters"], res["mixed"]["solve_s"],
               res["mixed"]["iters"], res["speedup_mixed_vs_fp64"]))
        out[name] = res
        del m
        torch.cuda.empty_cache()
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
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false")
    ap.add_argument("--sphere", choices=["off", "short", "full"], default="full",
                    help="icosahedral sphere advection: short = timed steps only, "
                         "full = one revolution with L2/Linf errors")
    ap.add_argument("--sphere-level", type=int, default=6)
    ap.add_argument("--sphere-steps", type=int, default=10000, help="RK4 steps per revolution")
    ap.add_argument("--sphere-nu", type=float, default=5e-5,
                    help="viscosity of the advected field (0: pure advection, which the "
                         "collocation scheme does not keep bounded on the sphere)")
    ap.add_argument("--ensemble", type=int, default=48,
                    help="members of the batched ensemble (0: off)")
    ap.add_argument("--ensemble-counts", default="8,12,16,24",
                    help="element counts per direction to draw from")
    ap.add_argument("--kernels", type=int, default=0, metavar="REPS",
                    help="only time the path's kernels REPS times each on the 32^3 p=7 boxes "
                         "(for profiler counter passes) and print their average times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (the kernels work on "
              "one GPU's share of the elements; several GPUs would split the element "
              "ranges and exchange the shared faces)" % args.gpus, file=sys.stderr)
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
    bg = ga.BoxGrid(3, p, [ne] * 3, device_maps=True)
    m = Mesh(ctx, bg)
    log("periodic box grid + mesh (GGFX maps on the device): %.1f s, ndof=%d"
        % (time.perf_counter() - t, bg.ndof))
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
    halo = bool(m.flags & ga._capi.GF_MESH_BOX_HALO)
    stage_ms = m.time_kernel("burgers_stage_box" if boxed else "burgers_stage", reps=20)
    stage_name = ("k_burgers_stage3b<8> (structured-box RK stage, K gathered in-kernel%s; "
                  "middle-stage form)" % (" through the LDS halo" if halo else "") if boxed else
                  "k_burgers_stage3p<8> (fused RK stage; middle-stage form)")
    stage_bytes = 5 * 8 * 3 * ndof  # middle-stage form: read uin,kprev,acc; write acc,k (per comp)
    stage_gbs = stage_bytes / (stage_ms * 1e-3) / 1e9
    step_bytes = 16 * 8 * 3 * ndof
    step_gbs = step_bytes / (el / args.steps) / 1e9
    out.update({"value": dofsteps, "unit": "DOF-steps/s", "ms_per_step": 1e3 * el / args.steps})
    traffic = (measured_traffic("k_burgers_stage3b<%d, true, 2, false, %s>"
                                % (p + 1, "true" if halo else "false")) if boxed else None)
    peak_meas = copy_peak(torch)
    log("measured copy bandwidth %.0f GB/s" % peak_meas)
    peaks_tf = {w: ctx.peak(w) for w in ("mfma_f32", "mfma_f64", "valu_f64")}
    log("measured peaks: " + ", ".join("%s %.1f TFLOP/s" % kv for kv in peaks_tf.items()))
    out["roofline"] = {"kernel": stage_name,
                       "bound": "hbm", "achieved": stage_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": stage_gbs / HBM_PEAK_GBS, "traffic": traffic,
                       "traffic_source": "profiles/pmc_32cube_p7_kernels.md" if traffic else None,
                       "avg_launch_ms": stage_ms, "alg_bytes_per_launch": stage_bytes,
                       "step_alg_bytes": step_bytes, "step_achieved_gbs": step_gbs,
                       "step_frac": step_gbs / HBM_PEAK_GBS,
                       "measured_peak": peak_meas,
                       "measured_peak_source": "torch copy_ of 134 MB (read + write), HIP events",
                       "measured_compute_peaks_tflops": peaks_tf,
                       "measured_compute_peaks_source": "gf_peak_probe: 4 waves/SIMD of "
                       "independent v_mfma_f32_32x32x2f32 / v_mfma_f64_16x16x4f64 / v_fma_f64 "
                       "chains on every CU, HIP events",
                       "frac_of_measured": stage_gbs / peak_meas,
                       "step_frac_of_measured": step_gbs / peak_meas}
    del us, m
    torch.cuda.empty_cache()

    # ---------------- Helmholtz CG -------------------------------------------
    if not args.no_cg:
        t = time.perf_counter()
        cg_grid = ga.BoxGrid(3, p, [ne] * 3, bdy=[ga.GF_BDY_DIRICHLET] * 6, device_maps=True)
        m = Mesh(ctx, cg_grid)
        log("Dirichlet box grid + mesh (GGFX maps on the device): %.1f s"
            % (time.perf_counter() - t))
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

    if not args.no_cg:
        out["cg_node_metric"] = node_metric_cg(ctx)
    if args.adjoint:
        out["adjoint"] = adjoint_run(ctx)
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
        out["cpu_baseline"] = cpu_baseline()
        log("CPU baseline %.3e DOF-steps/s" % out["cpu_baseline"]["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
