"""The C++ adapter (include/geoflow_hip.hpp) as reference-style code uses it.

tests/cpp/adapter_driver.cpp drives GHelmholtz, GGFX, GMass, GCG and the
Burgers RK stepper through the adapter and writes its device results to
files; here they are compared with the CPU oracle on the same inputs:
  * Helmholtz (p=0.5, q=1 as a field and q=2.5 as a one-entry q), its DSS
    sum and the mass apply: 1e-12 relative;
  * 10 identity-preconditioned CG iterations: 1e-11 relative;
  * 3 RK4 steps of nonlinear Burgers (nu=0.05, dt=1e-3): 1e-12 relative;
  * on a box whose GGFX groups are built on the device: GGrid::deriv and
    GGrid::dealias (pq = 8) at 1e-12 relative, GGrid::integrate to 1e-13;
    GMTK::grad, GGrid::compute_grefdiv, GAdvect::apply and
    GBurgers::dudt_impl at 1e-12 relative;
  * the Jacobi diagonal of the Dirichlet box's operator at 1e-12 relative;
  * on the icosahedral sphere (IcosGrid, p = 4, one bisection): mass,
    Helmholtz and 3 RK4 steps of advection at 1e-12 relative.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "adapter_driver")


def test_adapter_header_compiles_standalone():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-x", "c++", os.path.join(ROOT, "include", "geoflow_hip.hpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_driver_is_built_and_reports_errors_without_a_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert os.access(DRIVER, os.X_OK), "run __graft_entry__.build() first"
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "gf_ctx_create failed" in r.stderr


def _load(d, name):
    return np.fromfile(os.path.join(d, name), dtype=np.float64)


@pytest.mark.gpu
def test_adapter_matches_oracle(ctx, tmp_path):
    print("adapter_driver", DRIVER, flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    og = oracle.Grid.box(3, 4, [3, 2, 2], bdy=[oracle.BDY_DIRICHLET] * 6)
    x, y, z = og.get("x"), og.get("y"), og.get("z")
    u = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z) + 0.1 * x * y
    f = np.cos(2 * np.pi * x) * (1 + y) * (1 + z * z)
    q = np.ones(og.ndof)
    h = og.helmholtz(u, 0.5, q)
    assert relerr(_load(tmp_path, "helm.bin"), h) <= 1e-12
    h2 = og.helmholtz(u, 0.5, np.full(og.ndof, 2.5))
    assert relerr(_load(tmp_path, "helm_qscalar.bin"), h2) <= 1e-12
    assert relerr(_load(tmp_path, "helm_dss_sum.bin"), og.dss(h, "sum")) <= 1e-12
    b = og.mass_apply(f)
    assert np.array_equal(_load(tmp_path, "mass.bin"), b)
    xo, ret, it, _ = og.cg_solve(b, xb=np.zeros(og.ndof), p=0.5, q=q, maxit=10, tol=1e-300)
    st, iters = _load(tmp_path, "cg_meta.bin")
    assert (int(st), int(iters)) == (ret, it) == (1, 10)
    assert relerr(_load(tmp_path, "cg_x.bin"), xo) <= 1e-11

    gp = oracle.Grid.box(3, 5, [2, 2, 2])
    X, Y, Z = gp.get("x"), gp.get("y"), gp.get("z")
    us = [np.sin(2 * np.pi * (X + c * Y)) + 0.3 * np.cos(2 * np.pi * Z) for c in range(3)]
    for s in range(3):
        gp.burgers_step(us, 0.05, 1e-3, t=s * 1e-3)
    for c in range(3):
        assert relerr(_load(tmp_path, "burgers_u%d.bin" % c), us[c]) <= 1e-12

    # device-built GGFX maps; GGrid::deriv, dealias (pq = 8) and integrate
    gd = oracle.Grid.box(3, 5, [2, 3, 2])
    X, Y, Z = gd.get("x"), gd.get("y"), gd.get("z")
    v = np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + Z
    w = np.cos(2 * np.pi * (X + Z)) + 0.5 * Y
    assert relerr(_load(tmp_path, "deriv_x.bin"), gd.deriv(v, 1)) <= 1e-12
    assert relerr(_load(tmp_path, "dealias.bin"), gd.dealias(v, w, 8)) <= 1e-12
    assert abs(_load(tmp_path, "integrate.bin")[0] - gd.integrate(v)) <= 1e-13
    # GMTK::grad, GGrid::compute_grefdiv, GAdvect, GBurgers::dudt_impl; the
    # Jacobi diagonal of the Dirichlet box's operator
    wz = 0.2 * np.sin(2 * np.pi * Y)
    for c in range(3):
        assert relerr(_load(tmp_path, "grad_%d.bin" % c), gd.deriv(v, c + 1)) <= 1e-12
    assert relerr(_load(tmp_path, "grefdiv.bin"), gd.grefdiv([v, w, None], True)) <= 1e-12
    assert relerr(_load(tmp_path, "advect.bin"), gd.advect(v, [w, v, wz])) <= 1e-12
    ref = gd.burgers_rhs([v, w, wz], 0.05)
    for c in range(3):
        assert relerr(_load(tmp_path, "dudt_%d.bin" % c), ref[c]) <= 1e-12
    assert relerr(_load(tmp_path, "diag_inv.bin"), og.helmholtz_diag_inv(0.5, q)) <= 1e-12

    # IcosGrid: mass, Helmholtz and RK4 advection on the sphere
    oi = oracle.Grid.icos(4, 1, radius=1.0)
    X, Y, Z = oi.get("x"), oi.get("y"), oi.get("z")
    hb = X * Y + Z
    assert relerr(_load(tmp_path, "icos_mass.bin"), oi.mass_apply(hb)) <= 1e-12
    assert relerr(_load(tmp_path, "icos_helm.bin"),
                  oi.helmholtz(hb, 0.7, np.full(oi.ndof, 1.3))) <= 1e-12
    st = [hb.copy(), -Y, X.copy(), np.zeros(oi.ndof)]
    for s in range(3):
        oi.burgers_step(st, 1e-3, 1e-3, t=s * 1e-3, mode=1)
    assert relerr(_load(tmp_path, "icos_u.bin"), st[0]) <= 1e-12
