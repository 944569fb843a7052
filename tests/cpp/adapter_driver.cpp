// Exercises the C++ adapter (include/geoflow_hip.hpp) the way reference code
// drives GHelmholtz, GMass, GGFX, GCG, the Burgers RK stepper and GGrid's
// deriv, dealias and integrate, and writes
// the device results as raw little-endian doubles into the directory given as
// argv[1].  tests/test_cpp_adapter.py runs it on a GPU and compares the files
// with the CPU oracle.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "geoflow_hip.hpp"

using namespace geoflow_hip;

static std::string g_out;

static void save(const char *name, const std::vector<double> &v) {
  std::ofstream f(g_out + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUTDIR\n", argv[0]);
    return 2;
  }
  g_out = argv[1];
  const double pi = 3.14159265358979323846;
  try {
    Context ctx(0);
    State tmp;

    // Helmholtz (p = 0.5, q = 1) on a Dirichlet box, p = 4, 3x2x2 elements.
    BoxGrid g(ctx, 3, {4, 4, 4}, {3, 2, 2}, {0, 0, 0}, {1, 1, 1}, {0, 0, 0, 0, 0, 0});
    const size_t n = g.ndof();
    std::vector<double> hu(n), hf(n), ones(n, 1.0);
    for (size_t i = 0; i < n; i++) {
      const double x = g.xyz(0)[i], y = g.xyz(1)[i], z = g.xyz(2)[i];
      hu[i] = std::sin(pi * x) * std::sin(pi * y) * std::sin(pi * z) + 0.1 * x * y;
      hf[i] = std::cos(2 * pi * x) * (1 + y) * (1 + z * z);
    }
    Field u(ctx, hu), f(ctx, hf), q(ctx, ones), o = g.field();
    GHelmholtz H(g);
    H.set_Lap_scalar(0.5);
    H.set_mass_scalar(q);
    H.init();
    H.opVec_prod(u, tmp, o);
    save("helm.bin", o.download());
    GHelmholtz Hs(g);  // one-entry q: the scalar form of set_mass_scalar
    Hs.set_Lap_scalar(0.5);
    Hs.set_mass_scalar(Field(ctx, std::vector<double>{2.5}));
    Field o2 = g.field();
    Hs.opVec_prod(u, tmp, o2);
    save("helm_qscalar.bin", o2.download());
    GGFX gg(g);
    gg.doOp(o, GGFX::Sum);
    save("helm_dss_sum.bin", o.download());

    GMass M(g);
    Field b = g.field();
    M.opVec_prod(f, tmp, b);
    save("mass.bin", b.download());

    // GCG, identity preconditioner, 10 iterations (unreachable tolerance).
    GCG::Traits tr;
    tr.maxit = 10;
    tr.tol = 1e-300;
    tr.normtype = GF_NORM_INF;
    GCG cg(tr, g);
    Field x(ctx, std::vector<double>(n, 0.0)), xb(ctx, std::vector<double>(n, 0.0));
    const int st = cg.solve(H, b, xb, x);
    save("cg_x.bin", x.download());
    save("cg_meta.bin", {(double)st, (double)cg.iterations()});

    // Burgers RK4, nonlinear, periodic box p = 5, 2x2x2 elements, 3 steps.
    BoxGrid gp(ctx, 3, {5, 5, 5}, {2, 2, 2});
    const size_t np = gp.ndof();
    std::vector<Field> us;
    for (int c = 0; c < 3; c++) {
      std::vector<double> h(np);
      for (size_t i = 0; i < np; i++) {
        const double X = gp.xyz(0)[i], Y = gp.xyz(1)[i], Z = gp.xyz(2)[i];
        h[i] = std::sin(2 * pi * (X + c * Y)) + 0.3 * std::cos(2 * pi * Z);
      }
      us.emplace_back(ctx, h);
    }
    State su = {&us[0], &us[1], &us[2]};
    BurgersStepper rk(gp, 0, 0.05);
    rk.setOrder(4);
    for (int s = 0; s < 3; s++) rk.step(s * 1e-3, su, 1e-3);
    for (int c = 0; c < 3; c++) {
      char name[32];
      std::snprintf(name, sizeof name, "burgers_u%d.bin", c);
      save(name, us[c].download());
    }

    // A periodic box whose GGFX groups are built on the device: GGrid::deriv,
    // GGrid::dealias (order 8 through p = 5) and GGrid::integrate.
    BoxGrid gd(ctx, 3, {5, 5, 5}, {2, 3, 2}, {0, 0, 0}, {1, 1, 1},
               {GF_BDY_PERIODIC, GF_BDY_PERIODIC, GF_BDY_PERIODIC, GF_BDY_PERIODIC,
                GF_BDY_PERIODIC, GF_BDY_PERIODIC},
               0.001, true);
    const size_t nd = gd.ndof();
    std::vector<double> hv(nd), hw(nd);
    for (size_t i = 0; i < nd; i++) {
      const double X = gd.xyz(0)[i], Y = gd.xyz(1)[i], Z = gd.xyz(2)[i];
      hv[i] = std::sin(2 * pi * X) * std::cos(2 * pi * Y) + Z;
      hw[i] = std::cos(2 * pi * (X + Z)) + 0.5 * Y;
    }
    Field v(ctx, hv), w(ctx, hw), dv = gd.field(), pr = gd.field();
    gd.deriv(v, 1, tmp, dv);
    save("deriv_x.bin", dv.download());
    gd.dealias(v, w, pr, 8);
    save("dealias.bin", pr.download());
    save("integrate.bin", {gd.integrate(v, tmp)});

    // GMTK::grad, GGrid::compute_grefdiv, GAdvect and GBurgers::dudt_impl on
    // the same box; the Jacobi diagonal of the Dirichlet box's operator.
    std::vector<Field> gr;
    for (int c = 0; c < 3; c++) gr.emplace_back(ctx, nd);
    State sg = {&gr[0], &gr[1], &gr[2]};
    gd.grad(v, sg);
    for (int c = 0; c < 3; c++) {
      char name[32];
      std::snprintf(name, sizeof name, "grad_%d.bin", c);
      save(name, gr[c].download());
    }
    Field dvg = gd.field();
    gd.grefdiv({&v, &w, nullptr}, true, dvg);
    save("grefdiv.bin", dvg.download());
    std::vector<double> hz(nd);
    for (size_t i = 0; i < nd; i++) hz[i] = 0.2 * std::sin(2 * pi * gd.xyz(1)[i]);
    Field wz(ctx, hz);
    State vel = {&w, &v, &wz};
    Field po = gd.field();
    GAdvect adv(gd);
    adv.apply(v, vel, tmp, po);
    save("advect.bin", po.download());
    std::vector<Field> rhs;
    for (int c = 0; c < 3; c++) rhs.emplace_back(ctx, nd);
    State sr = {&rhs[0], &rhs[1], &rhs[2]};
    State su3 = {&v, &w, &wz};
    BurgersStepper rk3(gd, 0, 0.05);
    rk3.dudt(0.0, su3, sr);
    for (int c = 0; c < 3; c++) {
      char name[32];
      std::snprintf(name, sizeof name, "dudt_%d.bin", c);
      save(name, rhs[c].download());
    }
    Field dinv = g.field();
    H.diag_inv(dinv);
    save("diag_inv.bin", dinv.download());

    // The icosahedral sphere (p = 4, one bisection): GMass, GHelmholtz and
    // three RK4 steps of advection under solid-body rotation about z.
    IcosGrid gi(ctx, 4, 1);
    const size_t ni = gi.ndof();
    std::vector<double> hb(ni), c0(ni), c1(ni), c2(ni, 0.0);
    for (size_t i = 0; i < ni; i++) {
      const double X = gi.xyz(0)[i], Y = gi.xyz(1)[i], Z = gi.xyz(2)[i];
      hb[i] = X * Y + Z;
      c0[i] = -Y;
      c1[i] = X;
    }
    Field fb(ctx, hb), f0(ctx, c0), f1(ctx, c1), f2(ctx, c2);
    GMass Mi(gi);
    Field mb = gi.field();
    Mi.opVec_prod(fb, tmp, mb);
    save("icos_mass.bin", mb.download());
    GHelmholtz Hi(gi);
    Hi.set_Lap_scalar(0.7);
    Hi.set_mass_scalar(1.3);
    Field hi = gi.field();
    Hi.opVec_prod(fb, tmp, hi);
    save("icos_helm.bin", hi.download());
    State si = {&fb, &f0, &f1, &f2};
    BurgersStepper rki(gi, 1, 1e-3);
    for (int s = 0; s < 3; s++) rki.step(s * 1e-3, si, 1e-3);
    save("icos_u.bin", fb.download());
    ctx.synchronize();
  } catch (const Error &e) {
    std::fprintf(stderr, "adapter_driver: %s\n", e.what());
    return 1;
  }
  std::printf("adapter_driver: ok\n");
  return 0;
}
