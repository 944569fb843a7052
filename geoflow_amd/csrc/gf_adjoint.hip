// gf_adjoint.hip -- hand-written adjoint of the Burgers right-hand side and of
// the explicit RK step (nonlinear mode), and a checkpointed driver that
// returns the gradient of a final-time functional with respect to the
// initial state.
//
// Forward operators (GBurgers::dudt_impl, src/cdg/pdes/gburgers.ipp:203-257;
// GExRKStepper::step_b, src/cdg/pdes/gexrk_stepper.ipp:252-341):
//   F_c(u) = Minv * ( -nu L u_c - M * sum_j u_j * d_j u_c )
//   K_m = S(F(u_m)),  u_m = uin + h b_m K_{m-1},  acc = uin + sum_m h c_m K_m,
//   the final stage input and the output are Smoothed: S(...)
// with S the GGFX Smooth.  As matrices on local vectors, S (Sum/count over a
// coincidence group) and the element Laplacian L are symmetric, so the
// adjoint uses them unchanged; d_j^T = D_j^T scaled by the element metric,
// formed with the transposed tensor-product contractions.  Regular meshes
// (on embedded meshes the step also projects onto the sphere's tangent
// plane, whose transpose is not built).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gf_solve_common.hpp"

using namespace gfs;

namespace {

// y += a * x * b   (b may be NULL)
__global__ void k_axpy_mul(double *__restrict__ y, double a, const double *__restrict__ x,
                           const double *__restrict__ b, size_t n) {
  GRID_LOOP(g, n) y[g] += a * (b ? x[g] * b[g] : x[g]);
}
// o = a * b * c  (c may be NULL)
__global__ void k_mul3(double *__restrict__ o, const double *__restrict__ a,
                       const double *__restrict__ b, const double *__restrict__ c, size_t n) {
  GRID_LOOP(g, n) o[g] = c ? a[g] * b[g] * c[g] : a[g] * b[g];
}
// o = x + s * y
__global__ void k_xpsy(double *__restrict__ o, const double *__restrict__ x, double s,
                       const double *__restrict__ y, size_t n) {
  GRID_LOOP(g, n) o[g] = x[g] + s * y[g];
}
__global__ void k_copy_d(double *__restrict__ o, const double *__restrict__ x, size_t n) {
  GRID_LOOP(g, n) o[g] = x[g];
}

#define RCA(x)           \
  do {                   \
    int _rc = (x);       \
    if (_rc) return _rc; \
  } while (0)

int axpy_mul(gf_mesh *m, double *y, double a, const double *x, const double *b) {
  hipLaunchKernelGGL(k_axpy_mul, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, y, a,
                     x, b, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}
int mul3(gf_mesh *m, double *o, const double *a, const double *b, const double *c) {
  hipLaunchKernelGGL(k_mul3, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, o, a, b,
                     c, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}
int xpsy(gf_mesh *m, double *o, const double *x, double s, const double *y) {
  hipLaunchKernelGGL(k_xpsy, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, o, x, s, y,
                     m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}
int copyd(gf_mesh *m, double *o, const double *x) {
  hipLaunchKernelGGL(k_copy_d, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, o, x,
                     m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int nvel_of(gf_mesh *m) { return m->gtype == GF_GE_2DEMBEDDED ? 3 : m->dim; }

// out = d_j^T v (physical derivative transposed), j = 0..nvel-1
int deriv_T(gf_mesh *m, int j, const double *v, double *out, double *t1, double *t2) {
  if (m->gtype == GF_GE_REGULAR) return gf_grefderiv_t(m, j + 1, 1, v, out, m->edx);
  // embedded: sum_i D_i^T (dXi_i/dX_j * v)
  RCA(gf_mul_t(m, v, m->ndx[0 * 3 + j], t1));
  RCA(gf_grefderiv_t(m, 1, 1, t1, out, nullptr));
  RCA(gf_mul_t(m, v, m->ndx[1 * 3 + j], t1));
  RCA(gf_grefderiv_t(m, 2, 1, t1, t2, nullptr));
  return gf_axpy_t(m, 1.0, t2, out);
}

// lam[k] += (dF/du)^T w  at the state u (nonlinear Burgers, n_c = nvel).
int rhs_adjoint(gf_mesh *m, double nu, const double *const *u, const double *const *w,
                double *const *lam) {
  const size_t bytes = m->ndof * sizeof(double);
  double *z = (double *)gf_ws(m, WS_ADJ_TMP + 0, bytes);
  double *mz = (double *)gf_ws(m, WS_ADJ_TMP + 1, bytes);
  double *t = (double *)gf_ws(m, WS_ADJ_TMP + 2, bytes);
  double *t1 = (double *)gf_ws(m, WS_ADJ_TMP + 3, bytes);
  double *t2 = (double *)gf_ws(m, WS_ADJ_TMP + 4, bytes);
  double *t3 = (double *)gf_ws(m, WS_ADJ_TMP + 5, bytes);
  if (!z || !mz || !t || !t1 || !t2 || !t3) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  const int nv = nvel_of(m);
  for (int c = 0; c < nv; c++) {
    RCA(gf_mul_t(m, m->imass, w[c], z));  // z = Minv w_c
    // viscous part: -nu L^T z = -nu L z
    RCA(gf_helm_t(m, 1.0, nullptr, nullptr, 0.0, z, t));
    RCA(axpy_mul(m, lam[c], -nu, t, nullptr));
    RCA(gf_mul_t(m, m->mass, z, mz));  // M z
    for (int j = 0; j < nv; j++) {
      // velocity dependence: lam_j -= (d_j u_c) M z
      RCA(gf_deriv(m, j + 1, u[c], t));
      RCA(axpy_mul(m, lam[j], -1.0, t, mz));
      // transported quantity: lam_c -= d_j^T (u_j M z)
      RCA(gf_mul_t(m, u[j], mz, t3));
      RCA(deriv_T(m, j, t3, t, t1, t2));
      RCA(axpy_mul(m, lam[c], -1.0, t, nullptr));
    }
  }
  return GF_OK;
}

void butcher_tab(int order, double *beta /*4x4*/, double *c) {
  for (int i = 0; i < 16; i++) beta[i] = 0.0;
  for (int i = 0; i < 4; i++) c[i] = 0.0;
  // GButcherRK::computeCoeffs (src/cdg/pdes/gbutcherrk.ipp:80-127)
  if (order == 1) {
    c[0] = 1.0;
  } else if (order == 2) {
    c[0] = 0.5; c[1] = 0.5; beta[1 * 4 + 0] = 1.0;
  } else if (order == 3) {
    c[0] = 0.25; c[2] = 0.75; beta[1 * 4 + 0] = 1.0 / 3.0; beta[2 * 4 + 1] = 2.0 / 3.0;
  } else {
    c[0] = 1.0 / 6.0; c[1] = 1.0 / 3.0; c[2] = 1.0 / 3.0; c[3] = 1.0 / 6.0;
    beta[1 * 4 + 0] = 0.5; beta[2 * 4 + 1] = 0.5; beta[3 * 4 + 2] = 1.0;
  }
}

// Stage states of one step from uin (composed operators, step_b's DSS
// points): us[s][k] for s = 0..order-1 (us[order-1] is the smoothed final
// stage input).  K workspace: K[k].
int stage_states(gf_mesh *m, double nu, int order, double h, const double *const *uin,
                 std::vector<std::vector<double *>> &us, double *const *K) {
  const int nv = nvel_of(m);
  double beta[16], c[4];
  butcher_tab(order, beta, c);
  for (int k = 0; k < nv; k++) RCA(copyd(m, us[0][k], uin[k]));
  if (order == 1) {
    for (int k = 0; k < nv; k++) RCA(gf_dss_t(m, GF_DSS_SMOOTH, us[0][k]));
    return GF_OK;
  }
  for (int s = 1; s < order; s++) {
    const double *cu[3] = {us[s - 1][0], us[s - 1][1], nv > 2 ? us[s - 1][2] : nullptr};
    RCA(gf_burgers_rhs(m, 0, nu, nv, cu, K));
    for (int k = 0; k < nv; k++) {
      RCA(gf_dss_t(m, GF_DSS_SMOOTH, K[k]));
      RCA(xpsy(m, us[s][k], uin[k], h * beta[s * 4 + (s - 1)], K[k]));
      if (s == order - 1) RCA(gf_dss_t(m, GF_DSS_SMOOTH, us[s][k]));
    }
  }
  return GF_OK;
}

// lam: in dJ/du_{n+1}, out dJ/du_n, for one step from un.
int step_adjoint(gf_mesh *m, double nu, int order, double h, const double *const *un,
                 double *const *lam) {
  const int nv = nvel_of(m);
  const size_t bytes = m->ndof * sizeof(double);
  double beta[16], c[4];
  butcher_tab(order, beta, c);
  // workspace: stage states (order x nv), stage cotangents Kbar (order x nv),
  // accumulator cotangent, lam_in, scratch K / y
  std::vector<std::vector<double *>> us(order, std::vector<double *>(3, nullptr)),
      kb(order, std::vector<double *>(3, nullptr));
  double *abar[3] = {nullptr, nullptr, nullptr}, *lin[3] = {nullptr, nullptr, nullptr},
         *Kw[3] = {nullptr, nullptr, nullptr}, *y[3] = {nullptr, nullptr, nullptr};
  int slot = WS_ADJ_BEGIN;
  auto ws = [&](double *&p) {
    p = (double *)gf_ws(m, slot++, bytes);
    return p != nullptr;
  };
  bool ok = true;
  for (int s = 0; s < order; s++)
    for (int k = 0; k < nv; k++) ok = ok && ws(us[s][k]) && ws(kb[s][k]);
  for (int k = 0; k < nv; k++) ok = ok && ws(abar[k]) && ws(lin[k]) && ws(Kw[k]) && ws(y[k]);
  if (!ok) return gf_set_err(m->ctx, GF_ERR_ALLOC, "adjoint workspace");
  RCA(stage_states(m, nu, order, h, un, us, Kw));
  for (int s = 0; s < order; s++)
    for (int k = 0; k < nv; k++)
      GF_HIP(m->ctx, hipMemsetAsync(kb[s][k], 0, bytes, m->ctx->stream));
  // out = S(acc + h c_last K_last)
  for (int k = 0; k < nv; k++) {
    RCA(copyd(m, abar[k], lam[k]));
    RCA(gf_dss_t(m, GF_DSS_SMOOTH, abar[k]));               // abar = S lam
    RCA(axpy_mul(m, kb[order - 1][k], h * c[order - 1], abar[k], nullptr));
    RCA(copyd(m, lin[k], abar[k]));                          // acc = uin + ...
  }
  for (int s = 0; s < order - 1; s++)
    for (int k = 0; k < nv; k++) RCA(axpy_mul(m, kb[s][k], h * c[s], abar[k], nullptr));
  // stages in reverse
  for (int s = order - 1; s >= 0; s--) {
    // K_s = S(F(u_s)): ubar_s = J_F(u_s)^T S Kbar_s
    for (int k = 0; k < nv; k++) {
      RCA(gf_dss_t(m, GF_DSS_SMOOTH, kb[s][k]));
      GF_HIP(m->ctx, hipMemsetAsync(y[k], 0, bytes, m->ctx->stream));
    }
    const double *cu[3] = {us[s][0], us[s][1], nv > 2 ? us[s][2] : nullptr};
    const double *cw[3] = {kb[s][0], kb[s][1], nv > 2 ? kb[s][2] : nullptr};
    RCA(rhs_adjoint(m, nu, cu, cw, y));
    const bool final_in = s == order - 1;
    for (int k = 0; k < nv; k++) {
      if (final_in) RCA(gf_dss_t(m, GF_DSS_SMOOTH, y[k]));  // u_s = S(uin + h b K_{s-1})
      RCA(axpy_mul(m, lin[k], 1.0, y[k], nullptr));
      if (s >= 1) RCA(axpy_mul(m, kb[s - 1][k], h * beta[s * 4 + (s - 1)], y[k], nullptr));
    }
  }
  for (int k = 0; k < nv; k++) RCA(copyd(m, lam[k], lin[k]));
  return GF_OK;
}

}  // namespace

extern "C" {

int gf_burgers_rhs_adjoint(gf_mesh *m, double nu, int nstate, const double *const *u,
                           const double *const *w, double *const *lam) {
  if (!m || !u || !w || !lam) return GF_ERR_ARG;
  if (!m->Niso || m->gtype != GF_GE_REGULAR)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "adjoint: regular isotropic meshes");
  if (nstate != nvel_of(m)) return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: nonlinear mode needs nstate == velocity components");
  for (int k = 0; k < nstate; k++)
    if (!u[k] || !w[k] || !lam[k]) return GF_ERR_ARG;
  return rhs_adjoint(m, nu, u, w, lam);
}

int gf_burgers_step_adjoint(gf_mesh *m, double nu, int order, int nstate, double dt,
                            const double *const *un, double *const *lam) {
  if (!m || !un || !lam || order < 1 || order > 4) return GF_ERR_ARG;
  if (!m->Niso || m->gtype != GF_GE_REGULAR)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "adjoint: regular isotropic meshes");
  if (nstate != nvel_of(m)) return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: nonlinear mode needs nstate == velocity components");
  for (int k = 0; k < nstate; k++)
    if (!un[k] || !lam[k]) return GF_ERR_ARG;
  return step_adjoint(m, nu, order, dt, un, lam);
}

int gf_burgers_adjoint(gf_mesh *m, double nu, int order, int nstate, double dt, int nsteps,
                       int ckpt, const double *const *u0, double *const *lam, double *const *uT) {
  if (!m || !u0 || !lam || order < 1 || order > 4 || nsteps < 0 || ckpt < 1) return GF_ERR_ARG;
  if (!m->Niso || m->gtype != GF_GE_REGULAR)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "adjoint: regular isotropic meshes");
  if (nstate != nvel_of(m)) return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: nonlinear mode needs nstate == velocity components");
  const size_t bytes = m->ndof * sizeof(double);
  const int ncp = (nsteps + ckpt - 1) / ckpt;  // checkpoints at steps 0, ckpt, 2 ckpt, ...
  // checkpoint and segment storage: (ncp + ckpt + 1) states of nstate
  // fields, in the mesh workspace (gf_mesh_reserve can allocate it ahead)
  std::vector<double *> store((size_t)(ncp + ckpt + 1) * nstate, nullptr);
  for (size_t i = 0; i < store.size(); i++) {
    store[i] = (double *)gf_ws(m, WS_ADJ_CKPT + (int)i, bytes);
    if (!store[i]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "adjoint checkpoint workspace");
  }
  auto cp = [&](int i) { return &store[(size_t)i * nstate]; };
  auto seg = [&](int i) { return &store[(size_t)(ncp + i) * nstate]; };
  auto run = [&]() -> int {
    // forward sweep, checkpointing every ckpt steps
    double *cur[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < nstate; k++) {
      cur[k] = seg(0)[k];
      RCA(copyd(m, cur[k], u0[k]));
    }
    for (int s = 0; s < nsteps; s++) {
      if (s % ckpt == 0)
        for (int k = 0; k < nstate; k++) RCA(copyd(m, cp(s / ckpt)[k], cur[k]));
      RCA(gf_burgers_step(m, 0, nu, order, nstate, s * dt, dt, cur, 0));
    }
    if (uT)
      for (int k = 0; k < nstate; k++)
        if (uT[k]) RCA(copyd(m, uT[k], cur[k]));
    // reverse sweep: recompute each segment's states from its checkpoint
    for (int c = ncp - 1; c >= 0; c--) {
      const int s0 = c * ckpt, s1 = std::min(nsteps, s0 + ckpt);
      for (int k = 0; k < nstate; k++) RCA(copyd(m, seg(0)[k], cp(c)[k]));
      for (int s = s0 + 1; s < s1; s++) {  // state at step s from the one before
        for (int k = 0; k < nstate; k++) RCA(copyd(m, seg(s - s0)[k], seg(s - s0 - 1)[k]));
        double *st[3] = {seg(s - s0)[0], nstate > 1 ? seg(s - s0)[1] : nullptr,
                         nstate > 2 ? seg(s - s0)[2] : nullptr};
        RCA(gf_burgers_step(m, 0, nu, order, nstate, (s - 1) * dt, dt, st, 0));
      }
      for (int s = s1 - 1; s >= s0; s--) {
        const double *un[3] = {seg(s - s0)[0], nstate > 1 ? seg(s - s0)[1] : nullptr,
                               nstate > 2 ? seg(s - s0)[2] : nullptr};
        RCA(step_adjoint(m, nu, order, dt, un, lam));
      }
    }
    GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return GF_OK;
  };
  return run();
}

// gf_burgers_adjoint with the checkpoints written as GIO state files
// (GIO::write_state_impl layout, gio.ipp:588-690) on the forward sweep and
// read back on the reverse sweep: only one segment (ckpt + 1 states) stays in
// device memory.  Checkpoint c (the state at step c*ckpt) is the file
// "<prefix>.<c>.gio": ivers 0, the mesh's dim, element count, orders and
// element type, cycle = step, time = step*dt, multivar with one field per
// state component.
int gf_burgers_adjoint_gio(gf_mesh *m, double nu, int order, int nstate, double dt, int nsteps,
                           int ckpt, const double *const *u0, double *const *lam,
                           double *const *uT, const char *prefix) {
  if (!m || !u0 || !lam || !prefix || order < 1 || order > 4 || nsteps < 0 || ckpt < 1)
    return GF_ERR_ARG;
  if (!m->Niso || m->gtype != GF_GE_REGULAR)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "adjoint: regular isotropic meshes");
  if (nstate != nvel_of(m)) return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: nonlinear mode needs nstate == velocity components");
  const size_t n = m->ndof, bytes = n * sizeof(double);
  const int ncp = (nsteps + ckpt - 1) / ckpt;
  std::vector<double *> store((size_t)(ckpt + 1) * nstate, nullptr);
  for (size_t i = 0; i < store.size(); i++) {
    store[i] = (double *)gf_ws(m, WS_ADJ_CKPT + (int)i, bytes);
    if (!store[i]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "adjoint segment workspace");
  }
  auto seg = [&](int i) { return &store[(size_t)i * nstate]; };
  std::vector<double> host((size_t)nstate * n);
  std::vector<double *> hf(nstate);
  for (int k = 0; k < nstate; k++) hf[k] = host.data() + (size_t)k * n;
  const int porder[3] = {m->N[0] - 1, m->N[1] - 1, m->dim == 3 ? m->N[2] - 1 : 0};
  auto path = [&](int c) { return std::string(prefix) + "." + std::to_string(c) + ".gio"; };
  auto save = [&](int c, int step, double *const *st) -> int {
    for (int k = 0; k < nstate; k++)
      GF_HIP(m->ctx, hipMemcpyAsync(hf[k], st[k], bytes, hipMemcpyDeviceToHost, m->ctx->stream));
    GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    gf_gio_info info;
    memset(&info, 0, sizeof(info));
    info.ivers = 0;
    info.dim = m->dim;
    info.nelems = m->nelems;
    info.gtype = m->gtype;
    info.cycle = (uint64_t)step;
    info.time = step * dt;
    info.multivar = nstate > 1 ? 1 : 0;
    if (gf_gio_write(path(c).c_str(), &info, porder, nullptr, hf.data(), nstate, n) != GF_OK)
      return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: cannot write a GIO checkpoint");
    return GF_OK;
  };
  auto load = [&](int c, int step, double *const *st) -> int {
    gf_gio_info info;
    memset(&info, 0, sizeof(info));
    if (gf_gio_read(path(c).c_str(), &info, nullptr, 0, nullptr, 0, hf.data(), nstate, n) != GF_OK)
      return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: cannot read a GIO checkpoint");
    if (info.cycle != (uint64_t)step || info.nelems != m->nelems)
      return gf_set_err(m->ctx, GF_ERR_ARG, "adjoint: GIO checkpoint of another run");
    for (int k = 0; k < nstate; k++)
      GF_HIP(m->ctx, hipMemcpyAsync(st[k], hf[k], bytes, hipMemcpyHostToDevice, m->ctx->stream));
    return GF_OK;
  };
  double *cur[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < nstate; k++) {
    cur[k] = seg(0)[k];
    RCA(copyd(m, cur[k], u0[k]));
  }
  for (int s = 0; s < nsteps; s++) {
    if (s % ckpt == 0) RCA(save(s / ckpt, s, cur));
    RCA(gf_burgers_step(m, 0, nu, order, nstate, s * dt, dt, cur, 0));
  }
  if (uT)
    for (int k = 0; k < nstate; k++)
      if (uT[k]) RCA(copyd(m, uT[k], cur[k]));
  for (int c = ncp - 1; c >= 0; c--) {
    const int s0 = c * ckpt, s1 = std::min(nsteps, s0 + ckpt);
    RCA(load(c, s0, seg(0)));
    for (int s = s0 + 1; s < s1; s++) {
      for (int k = 0; k < nstate; k++) RCA(copyd(m, seg(s - s0)[k], seg(s - s0 - 1)[k]));
      double *st[3] = {seg(s - s0)[0], nstate > 1 ? seg(s - s0)[1] : nullptr,
                       nstate > 2 ? seg(s - s0)[2] : nullptr};
      RCA(gf_burgers_step(m, 0, nu, order, nstate, (s - 1) * dt, dt, st, 0));
    }
    for (int s = s1 - 1; s >= s0; s--) {
      const double *un[3] = {seg(s - s0)[0], nstate > 1 ? seg(s - s0)[1] : nullptr,
                             nstate > 2 ? seg(s - s0)[2] : nullptr};
      RCA(step_adjoint(m, nu, order, dt, un, lam));
    }
  }
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return GF_OK;
}

}  // extern "C"
