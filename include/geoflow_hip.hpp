// geoflow_hip.hpp -- header-only C++17 adapter over the C ABI in geoflow_hip.h.
//
// Mirrors the reference's operator classes for the SEM hot path, with the
// same names and argument meaning, so that code written against
//   GMass::opVec_prod          (src/cdg/sem/gmass.hpp:37-39)
//   GHelmholtz::set_Lap_scalar / set_mass_scalar / init / opVec_prod
//                              (src/cdg/sem/ghelmholtz.hpp:58-64)
//   GGFX::init / doOp          (src/cdg/comm/ggfx.hpp:87-90)
//   GCG::solve(A, b, xb, x)    (src/cdg/solvers/gcg.hpp:60-69,
//                               src/pdeint/lin_solver_base.hpp:38-44)
//   GExRKStepper::setOrder / step with the GBurgers right-hand side
//                              (src/cdg/pdes/gexrk_stepper.hpp:40-47,
//                               src/cdg/pdes/gburgers.hpp:118-143)
// can be pointed at device-resident state.  Fields are DeviceVector<double>
// objects holding element-contiguous local vectors (the GTVector layout).
// Errors become geoflow_hip::Error exceptions carrying the C status code;
// GCG's convergence codes (GCGERR_*) are returned, not thrown, as in the
// reference.
#ifndef GEOFLOW_HIP_HPP
#define GEOFLOW_HIP_HPP

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "geoflow_hip.h"

namespace geoflow_hip {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int st, gf_ctx *ctx, const char *what) {
  if (st == GF_OK) return;
  std::string m = std::string(what) + " failed (status " + std::to_string(st) + ")";
  if (ctx) m += ": " + std::string(gf_ctx_last_error(ctx));
  throw Error(st, m);
}

// One device and one HIP stream; all operator calls are enqueued on it.
class Context {
 public:
  explicit Context(int device = 0, void *stream = nullptr) {
    check(gf_ctx_create(device, stream, &h_), nullptr, "gf_ctx_create");
  }
  ~Context() {
    if (h_) gf_ctx_destroy(h_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  void synchronize() { check(gf_ctx_synchronize(h_), h_, "gf_ctx_synchronize"); }
  gf_ctx *handle() const { return h_; }

 private:
  gf_ctx *h_ = nullptr;
};

// Owning device array (the role GTVector<T> plays for one state component).
template <typename T>
class DeviceVector {
 public:
  DeviceVector() = default;
  DeviceVector(Context &ctx, size_t n) : ctx_(ctx.handle()), n_(n) {
    void *p = nullptr;
    check(gf_malloc(ctx_, n * sizeof(T), &p), ctx_, "gf_malloc");
    p_ = static_cast<T *>(p);
  }
  DeviceVector(Context &ctx, const std::vector<T> &h) : DeviceVector(ctx, h.size()) { upload(h); }
  ~DeviceVector() { release(); }
  DeviceVector(DeviceVector &&o) noexcept { *this = std::move(o); }
  DeviceVector &operator=(DeviceVector &&o) noexcept {
    if (this != &o) {
      release();
      ctx_ = o.ctx_;
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;

  T *data() { return p_; }
  const T *data() const { return p_; }
  size_t size() const { return n_; }
  void upload(const std::vector<T> &h) {
    if (h.size() != n_) throw Error(GF_ERR_ARG, "DeviceVector::upload: size mismatch");
    check(gf_memcpy_h2d(ctx_, p_, h.data(), n_ * sizeof(T)), ctx_, "gf_memcpy_h2d");
  }
  std::vector<T> download() const {
    std::vector<T> h(n_);
    check(gf_memcpy_d2h(ctx_, h.data(), p_, n_ * sizeof(T)), ctx_, "gf_memcpy_d2h");
    return h;
  }

 private:
  void release() {
    if (p_) gf_free(ctx_, p_);
    p_ = nullptr;
  }
  gf_ctx *ctx_ = nullptr;
  T *p_ = nullptr;
  size_t n_ = 0;
};

using Field = DeviceVector<double>;
using State = std::vector<Field *>;

// GGrid (src/cdg/grid/ggrid.hpp): a grid's device mesh and the operations of
// GGrid that every grid type shares; BoxGrid and IcosGrid build it.
class Grid {
 public:
  virtual ~Grid() {
    if (mesh_) gf_mesh_destroy(mesh_);
  }
  Grid(const Grid &) = delete;
  Grid &operator=(const Grid &) = delete;

  Context &context() { return ctx_; }
  gf_mesh *mesh() const { return mesh_; }
  int dim() const { return dim_; }
  size_t ndof() const { return ndof_; }
  size_t nelems() const { return nelems_; }
  double minnodedist() const { return minnodedist_; }
  const std::vector<double> &xyz(int d) const { return xyz_[d]; }
  Field field() { return Field(ctx_, ndof_); }

  // GGrid::deriv (src/cdg/grid/ggrid.ipp:1163-1191): du = d u / dx_idir
  // (idir 1..dim, 1..3 on embedded grids).
  void deriv(const Field &u, int idir, State & /*utmp*/, Field &du) {
    check(gf_deriv(mesh_, idir, u.data(), du.data()), ctx_.handle(), "gf_deriv");
  }
  // GGrid::compute_grefderiv (src/cdg/grid/ggrid.ipp:1954-2265): reference-
  // space derivative along idir, D^T when dotrans.
  void grefderiv(const Field &u, int idir, bool dotrans, Field &du) {
    check(gf_grefderiv(mesh_, idir, dotrans ? 1 : 0, u.data(), du.data()), ctx_.handle(),
          "gf_grefderiv");
  }
  // GGrid::compute_grefdiv (src/cdg/grid/ggrid.ipp:1847-1927): divu =
  // sum_j D_j^T u_j (dotrans) or D_j u_j, reference space; entries of u may
  // be null (skipped).
  void grefdiv(const std::vector<const Field *> &u, bool dotrans, Field &divu) {
    std::vector<const double *> up(dim_, nullptr);
    for (size_t k = 0; k < u.size() && k < (size_t)dim_; k++) up[k] = u[k] ? u[k]->data() : nullptr;
    check(gf_grefdiv(mesh_, dotrans ? 1 : 0, up.data(), divu.data()), ctx_.handle(),
          "gf_grefdiv");
  }
  // GMTK::grad (src/cdg/blas/gmtk.ipp:2883-2892): du[k] = d u / d x_k for
  // k < dim (3 components on embedded grids).
  void grad(const Field &u, State &du) {
    std::vector<double *> dp(du.size());
    for (size_t k = 0; k < du.size(); k++) dp[k] = du[k]->data();
    check(gf_grad(mesh_, u.data(), dp.data()), ctx_.handle(), "gf_grad");
  }
  // GGrid::integrate (src/cdg/grid/ggrid.ipp:1077-1148).
  double integrate(const Field &u, State & /*utmp*/) {
    double r = 0.0;
    check(gf_integrate(mesh_, u.data(), &r), ctx_.handle(), "gf_integrate");
    return r;
  }

 protected:
  Grid(Context &ctx, int dim) : ctx_(ctx), dim_(dim) {}
  Context &ctx_;
  int dim_;
  size_t ndof_ = 0, nelems_ = 0;
  double minnodedist_ = 0;
  std::array<std::vector<double>, 3> xyz_;
  gf_mesh *mesh_ = nullptr;
};

// GGridBox for one rank: nodes, regular-element metrics, boundary mask and the
// GGFX coincidence groups, built on the host; operator data uploaded to the
// device as one mesh (src/cdg/grid/ggrid_box.ipp:254-463, 753-802, 862-978;
// src/cdg/comm/ggfx.hpp:167-342).  Periodic directions use GF_BDY_PERIODIC.
class BoxGrid : public Grid {
 public:
  BoxGrid(Context &ctx, int dim, std::array<int, 3> p, std::array<int, 3> ne,
          std::array<double, 3> P0 = {0, 0, 0}, std::array<double, 3> P1 = {1, 1, 1},
          std::array<int, 6> bdy = {GF_BDY_PERIODIC, GF_BDY_PERIODIC, GF_BDY_PERIODIC,
                                    GF_BDY_PERIODIC, GF_BDY_PERIODIC, GF_BDY_PERIODIC},
          double tol_fac = 0.001, bool device_maps = false)
      : Grid(ctx, dim) {
    if (dim != 2 && dim != 3) throw Error(GF_ERR_ARG, "BoxGrid: dim must be 2 or 3");
    size_t nn = 1, nel = 1;
    for (int d = 0; d < dim; d++) {
      nn *= (size_t)(p[d] + 1);
      nel *= (size_t)ne[d];
    }
    nelems_ = nel;
    ndof_ = nn * nel;
    for (int d = 0; d < dim; d++) {
      xyz_[d].resize(ndof_);
      xyzp_[d].resize(ndof_);
    }
    mask_.resize(ndof_);
    edx_.resize(nel * dim);
    ejac_.resize(nel);
    double *xp[3] = {nullptr, nullptr, nullptr}, *xpp[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < dim; d++) {
      xp[d] = xyz_[d].data();
      xpp[d] = xyzp_[d].data();
    }
    check(gf_box_grid(dim, p.data(), ne.data(), P0.data(), P1.data(), bdy.data(), xp, xpp,
                      mask_.data(), edx_.data(), ejac_.data(), &minnodedist_),
          nullptr, "gf_box_grid");
    const double *cx[3] = {xpp[0], xpp[1], xpp[2]};
    // device_maps: the GGFX groups are formed on the device from the box
    // lattice at mesh creation (3-D, every direction periodic or none)
    if (!device_maps) {
      offs_.resize(ndof_ + 1);
      size_t nm = 0;
      const double tol = tol_fac * minnodedist_;
      check(gf_ggfx_build(ndof_, dim, cx, tol, (size_t)1 << dim, offs_.data(), nullptr, &nm),
            nullptr, "gf_ggfx_build");
      members_.resize(nm);
      check(gf_ggfx_build(ndof_, dim, cx, tol, (size_t)1 << dim, offs_.data(), members_.data(),
                          &nm),
            nullptr, "gf_ggfx_build");
    }
    for (int d = 0; d < dim; d++) {
      std::vector<double> x(p[d] + 1), w(p[d] + 1), D((p[d] + 1) * (p[d] + 1));
      check(gf_gll_basis(p[d], x.data(), w.data(), D.data()), nullptr, "gf_gll_basis");
      W_[d] = std::move(w);
      D_[d] = std::move(D);
    }
    gf_mesh_desc md{};
    md.dim = dim;
    md.gtype = GF_GE_REGULAR;
    md.nelems = nel;
    for (int d = 0; d < 3; d++) md.N[d] = d < dim ? p[d] + 1 : 1;
    for (int d = 0; d < dim; d++) {
      md.D[d] = D_[d].data();
      md.W[d] = W_[d].data();
      md.xyz[d] = xyz_[d].data();
    }
    md.elem_dxidx = edx_.data();
    md.elem_jac = ejac_.data();
    md.mask = mask_.data();
    md.ggfx_offs = device_maps ? nullptr : offs_.data();
    md.ggfx_members = device_maps ? nullptr : members_.data();
    md.box_device_maps = device_maps ? 1 : 0;
    if (dim == 3)
      for (int d = 0; d < 3; d++) {
        md.box_ne[d] = ne[d];
        md.box_periodic[d] = bdy[2 * d] == GF_BDY_PERIODIC && bdy[2 * d + 1] == GF_BDY_PERIODIC;
      }
    check(gf_mesh_create(ctx.handle(), &md, &mesh_), ctx.handle(), "gf_mesh_create");
  }
  const std::vector<double> &mask() const { return mask_; }

  // GGrid::dealias (src/cdg/grid/ggrid.ipp:2424-2503) through the order-pq
  // GLL grid (the reference's qdealias_order).
  void dealias(const Field &v1, const Field &v2, Field &prod, int pq) {
    check(gf_dealias(mesh_, pq, v1.data(), v2.data(), prod.data()), ctx_.handle(), "gf_dealias");
  }
 private:
  std::array<std::vector<double>, 3> xyzp_, W_, D_;
  std::vector<double> mask_, edx_, ejac_;
  std::vector<int64_t> offs_, members_;
};

// GGridIcos (2-D) for one rank (src/cdg/grid/ggrid_icos.ipp:139-513): the
// base icosahedron refined (refine 0: GICOS_BISECTION with ilevel
// bisections; 1: GICOS_LAGRANGIAN with ilevel rows per base triangle), three
// GE_2DEMBEDDED quads per triangle on the sphere of the given radius, the
// embedded metrics of GElem_base (src/cdg/sem/gelem_base.cpp:877-955,
// 1066-1170, 1249-1300) and the GGFX groups (at most 6 members,
// GGridIcos::max_duplicates).
class IcosGrid : public Grid {
 public:
  IcosGrid(Context &ctx, int p, int ilevel, double radius = 1.0, int refine = 0,
           double tol_fac = 0.001)
      : Grid(ctx, 2) {
    size_t nel = 0;
    check(gf_icos_grid(p, ilevel, refine, radius, &nel, nullptr, nullptr, nullptr, nullptr),
          nullptr, "gf_icos_grid");
    nelems_ = nel;
    ndof_ = nel * (size_t)(p + 1) * (size_t)(p + 1);
    for (int d = 0; d < 3; d++) xyz_[d].resize(ndof_);
    for (auto &v : dxidx_) v.resize(ndof_);
    jac_.resize(ndof_);
    double *xp[3] = {xyz_[0].data(), xyz_[1].data(), xyz_[2].data()};
    double *dp[6];
    for (int k = 0; k < 6; k++) dp[k] = dxidx_[k].data();
    check(gf_icos_grid(p, ilevel, refine, radius, &nel, xp, dp, jac_.data(), &minnodedist_),
          nullptr, "gf_icos_grid");
    const double *cx[3] = {xp[0], xp[1], xp[2]};
    offs_.resize(ndof_ + 1);
    size_t nm = 0;
    const double tol = tol_fac * minnodedist_;
    check(gf_ggfx_build(ndof_, 3, cx, tol, 6, offs_.data(), nullptr, &nm), nullptr,
          "gf_ggfx_build");
    members_.resize(nm);
    check(gf_ggfx_build(ndof_, 3, cx, tol, 6, offs_.data(), members_.data(), &nm), nullptr,
          "gf_ggfx_build");
    W_.resize(p + 1);
    D_.resize((size_t)(p + 1) * (p + 1));
    std::vector<double> xi(p + 1);
    check(gf_gll_basis(p, xi.data(), W_.data(), D_.data()), nullptr, "gf_gll_basis");
    gf_mesh_desc md{};
    md.dim = 2;
    md.gtype = GF_GE_2DEMBEDDED;
    md.nelems = nel;
    md.N[0] = md.N[1] = p + 1;
    md.N[2] = 1;
    for (int d = 0; d < 2; d++) {
      md.D[d] = D_.data();
      md.W[d] = W_.data();
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 3; j++) md.node_dxidx[i * 3 + j] = dxidx_[i * 3 + j].data();
    md.node_jac = jac_.data();
    for (int d = 0; d < 3; d++) md.xyz[d] = xyz_[d].data();
    md.ggfx_offs = offs_.data();
    md.ggfx_members = members_.data();
    check(gf_mesh_create(ctx.handle(), &md, &mesh_), ctx.handle(), "gf_mesh_create");
  }

 private:
  std::array<std::vector<double>, 6> dxidx_;
  std::vector<double> jac_, W_, D_;
  std::vector<int64_t> offs_, members_;
};

// GMass (src/cdg/sem/gmass.hpp:37-44): out = M in, unassembled; inverse
// applies 1/DSS-Smooth(M).
class GMass {
 public:
  explicit GMass(Grid &g, bool inverse = false) : g_(g), inv_(inverse) {}
  void opVec_prod(const Field &in, State & /*utmp*/, Field &out) {
    gf_ctx *c = g_.context().handle();
    if (inv_)
      check(gf_imass_apply(g_.mesh(), in.data(), out.data()), c, "gf_imass_apply");
    else
      check(gf_mass_apply(g_.mesh(), in.data(), out.data()), c, "gf_mass_apply");
  }

 private:
  Grid &g_;
  bool inv_;
};

// GHelmholtz (src/cdg/sem/ghelmholtz.hpp:58-64): uo = p L u + q M u.  As in
// the reference, the mass term is present only after set_mass_scalar, the
// Laplacian factor defaults to 1, and a p or q with fewer entries than the
// grid has nodes is the scalar p[0] or q[0] (ghelmholtz.ipp:184-186).
class GHelmholtz {
 public:
  explicit GHelmholtz(Grid &g) : g_(g) {}
  void set_Lap_scalar(const Field &p) {
    if (p.size() < g_.ndof()) {
      pfield_ = nullptr;
      pscalar_ = p.download()[0];
    } else {
      pfield_ = p.data();
    }
  }
  void set_Lap_scalar(double p) {
    pfield_ = nullptr;
    pscalar_ = p;
  }
  void set_mass_scalar(const Field &q) {
    if (q.size() < g_.ndof()) {
      qfield_ = nullptr;
      qscalar_ = q.download()[0];
    } else {
      qfield_ = q.data();
      qscalar_ = 0.0;
    }
  }
  void set_mass_scalar(double q) {
    qfield_ = nullptr;
    qscalar_ = q;
  }
  void init() {}
  void use_metric(bool) {}
  void opVec_prod(const Field &in, State & /*utmp*/, Field &out) {
    check(gf_helmholtz_apply(g_.mesh(), pscalar_, pfield_, qscalar_, qfield_, in.data(),
                             out.data()),
          g_.context().handle(), "gf_helmholtz_apply");
  }
  // Inverse of the assembled diagonal of p L + q M with scalar p and the
  // current mass term (the Jacobi preconditioner GCG's precond = 1 uses; no
  // reference counterpart: GCG ships no preconditioner).
  void diag_inv(Field &dinv) {
    check(gf_helmholtz_diag_inv(g_.mesh(), pscalar_, qscalar_, qfield_, dinv.data()),
          g_.context().handle(), "gf_helmholtz_diag_inv");
  }
  Grid &grid() { return g_; }
  double p_scalar() const { return pscalar_; }
  const double *p_field() const { return pfield_; }
  const double *q_field() const { return qfield_; }
  double q_scalar() const { return qscalar_; }

 private:
  Grid &g_;
  double pscalar_ = 1.0;
  const double *pfield_ = nullptr;
  const double *qfield_ = nullptr;
  double qscalar_ = 0.0;  // no mass term until set_mass_scalar
};

// GGFX (src/cdg/comm/ggfx.hpp:87-90): the groups are built with the grid;
// doOp performs the in-place reduction.
class GGFX {
 public:
  enum ReductionOp { Sum = GF_DSS_SUM, Smooth = GF_DSS_SMOOTH, Max = GF_DSS_MAX, Min = GF_DSS_MIN };
  explicit GGFX(Grid &g) : g_(g) {}
  bool doOp(Field &u, ReductionOp op) {
    check(gf_dss(g_.mesh(), (int)op, u.data()), g_.context().handle(), "gf_dss");
    return true;
  }
  void get_imult(Field &imult) {
    check(gf_imult(g_.mesh(), imult.data()), g_.context().handle(), "gf_imult");
  }

 private:
  Grid &g_;
};

// GCG (src/cdg/solvers/gcg.hpp:29-75) with LinSolverBase::Traits
// (src/pdeint/lin_solver_base.hpp:38-44), plus the Jacobi and mixed-precision
// options of this implementation.  solve() returns the GCGERR code.
class GCG {
 public:
  struct Traits {
    bool bbdycond = false;
    int maxit = 512;
    int normtype = GF_NORM_INF;
    double tol = 1e-6;
    int precond = 0;      // 0 identity (reference), 1 Jacobi
    int precision = 0;    // 0 FP64, 1 FP32 recurrences with FP64 reliable updates
    double inner_tol = 0; // 0 = library default
  };
  GCG(Traits &traits, Grid &grid) : traits_(traits), g_(grid) {}
  int solve(GHelmholtz &A, const Field &b, Field &x) { return run(A, b, nullptr, x); }
  int solve(GHelmholtz &A, const Field &b, const Field &xb, Field &x) {
    return run(A, b, xb.data(), x);
  }
  int iterations() const { return res_.iters; }
  int outer_iterations() const { return res_.outer; }
  double get_resid_max() const { return res_.rnorm; }
  const std::vector<double> &get_residuals() const { return resid_; }

 private:
  int run(GHelmholtz &A, const Field &b, const double *xb, Field &x) {
    gf_cg_params prm{};
    prm.maxit = traits_.maxit;
    prm.tol = traits_.tol;
    prm.normtype = traits_.normtype;
    prm.precond = traits_.precond;
    prm.precision = traits_.precision;
    prm.inner_tol = traits_.inner_tol;
    prm.p_scalar = A.p_scalar();
    prm.p_field = A.p_field();
    prm.q_field = A.q_field();
    prm.q_scalar = A.q_scalar();
    resid_.assign(traits_.maxit > 0 ? traits_.maxit : 1, 0.0);
    prm.resid = resid_.data();
    int st = gf_cg_solve(g_.mesh(), &prm, b.data(), xb, x.data(), &res_);
    if (st >= GF_ERR_ARG) check(st, g_.context().handle(), "gf_cg_solve");
    resid_.resize(res_.iters < (int)resid_.size() ? res_.iters : resid_.size());
    return st;
  }
  Traits traits_;
  Grid &g_;
  gf_cg_result res_{};
  std::vector<double> resid_;
};

// GAdvect (src/cdg/sem/gadvect.hpp): po = M (u . grad p), unassembled, for
// the velocity components u (apply with ivec = -2, gadvect.ipp:57-99).
class GAdvect {
 public:
  explicit GAdvect(Grid &g) : g_(g) {}
  void apply(const Field &p, const State &u, State & /*utmp*/, Field &po) {
    std::vector<const double *> up(u.size());
    for (size_t k = 0; k < u.size(); k++) up[k] = u[k]->data();
    check(gf_advect(g_.mesh(), p.data(), (int)u.size(), up.data(), po.data()),
          g_.context().handle(), "gf_advect");
  }

 private:
  Grid &g_;
};

// GExRKStepper (src/cdg/pdes/gexrk_stepper.hpp:20-60) bound to the GBurgers
// right-hand side (src/cdg/pdes/gburgers.ipp:202-257): mode 0 nonlinear,
// 1 pure advection of u[0] by u[1..], 2 heat.  step() advances u in place by
// one explicit RK step of order 1..4 with GGFX Smooth after each stage.
class BurgersStepper {
 public:
  BurgersStepper(Grid &g, int mode, double nu) : g_(g), mode_(mode), nu_(nu) {}
  void setOrder(int order, int /*nstage*/ = 0) { order_ = order; }
  void step(double t, State &u, double dt, bool check_finite = false) {
    std::vector<double *> up(u.size());
    for (size_t k = 0; k < u.size(); k++) up[k] = u[k]->data();
    check(gf_burgers_step(g_.mesh(), mode_, nu_, order_, (int)u.size(), t, dt, up.data(),
                          check_finite ? 1 : 0),
          g_.context().handle(), "gf_burgers_step");
  }
  void steps(double t, State &u, double dt, int nsteps) {
    std::vector<double *> up(u.size());
    for (size_t k = 0; k < u.size(); k++) up[k] = u[k]->data();
    check(gf_burgers_steps(g_.mesh(), mode_, nu_, order_, (int)u.size(), t, dt, nsteps, up.data()),
          g_.context().handle(), "gf_burgers_steps");
  }
  // GBurgers::dudt_impl (src/cdg/pdes/gburgers.ipp:202-257): the right-hand
  // side the steps integrate, dudt[k] for the evolved components.
  void dudt(double /*t*/, const State &u, State &dudt) {
    std::vector<const double *> up(u.size());
    for (size_t k = 0; k < u.size(); k++) up[k] = u[k]->data();
    std::vector<double *> dp(dudt.size());
    for (size_t k = 0; k < dudt.size(); k++) dp[k] = dudt[k]->data();
    check(gf_burgers_rhs(g_.mesh(), mode_, nu_, (int)u.size(), up.data(), dp.data()),
          g_.context().handle(), "gf_burgers_rhs");
  }

 private:
  Grid &g_;
  int mode_;
  double nu_;
  int order_ = 4;
};

}  // namespace geoflow_hip

#endif
