// gf_oracle.cpp -- CPU restatement of the GeoFLOW SEM operator path.
//
// TEST INFRASTRUCTURE ONLY (see gf_oracle.h).  Compiled with -ffp-contract=off
// so that every sum and product rounds exactly as written, in the order the
// reference writes it.  Citations are file:line in the GeoFLOW source tree.
#include "gf_oracle.h"

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <set>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <functional>
#include <omp.h>
#endif

namespace {

typedef long double LD;

// ---------------------------------------------------------------------------
// GLL basis.  Follows GLLBasis<long double,double>::computeJacobi,
// computeNodes, computeWeights, computeDerivMatrix
// (src/cdg/sem/gllbasis.ipp:410-501, 514-531, 579-607, 618-667).
// ---------------------------------------------------------------------------
void jacobi(int N, LD alpha, LD beta, LD &poly, LD &pder, LD &polym1, LD &pderm1,
            LD &polym2, LD &pderm2, LD x) {
  // gllbasis.ipp:618-667
  LD apb = alpha + beta, rv = 1.0L + alpha;
  LD polylist = 0, pderlist = 0, psave = 0, pdsave = 0;
  poly = 1.0L;
  pder = 0.0L;
  if (N == 0) return;
  polylist = poly;
  pderlist = pder;
  poly = rv * x;
  pder = rv;
  if (N == 1) return;
  for (int k = 2; k <= N; k++) {
    LD fk = (LD)k;
    LD a1 = 2.0L * fk * (fk + apb) * (2.0L * fk + apb - 2.0L);
    LD a2 = (2.0L * fk + apb - 1.0L) * (alpha * alpha - beta * beta);
    LD b3 = 2.0L * fk + apb - 2.0L;
    LD a3 = b3 * (b3 + 1.0L) * (b3 + 2.0L);
    LD a4 = 2.0L * (fk + alpha - 1.0L) * (fk + beta - 1.0L) * (2.0L * fk + apb);
    LD polyn = ((a2 + a3 * x) * poly - a4 * polylist) / a1;
    LD pdern = ((a2 + a3 * x) * pder - a4 * pderlist + a3 * poly) / a1;
    psave = polylist;
    pdsave = pderlist;
    polylist = poly;
    poly = polyn;
    pderlist = pder;
    pder = pdern;
  }
  polym1 = polylist;
  pderm1 = pderlist;
  polym2 = psave;
  pderm2 = pdsave;
}

struct Basis {
  int p = 0, N = 0;
  std::vector<double> xi, w, D, DT;  // D(l,j) column-major: D[l + j*N]
  std::vector<LD> xl, Pnl;           // long-double nodes and P_p(node) (xiNodes_, Pn_)
};

bool gll(int Np, Basis &b) {
  if (Np < 1) return false;
  const int kstop = 128;
  const LD eps = 1.0e-8;  // eps_(1.0e-8), a double literal (gllbasis.hpp:43)
  const LD ttiny = 100.0L * std::numeric_limits<LD>::epsilon();
  const LD ximin = -1.0L, ximax = 1.0L;
  const LD alpha = 0, beta = 0;
  int np1 = Np + 1;
  std::vector<LD> x(np1), Pn(np1), dPn(np1), w(np1);
  LD pnp1p, pdnp1p, pnp, pdnp, pnm1p, pdnm1, pnp1m, pdnp1m, pnm, pdnm, pnm1m;
  jacobi(np1, alpha, beta, pnp1p, pdnp1p, pnp, pdnp, pnm1p, pdnm1, ximax);
  jacobi(np1, alpha, beta, pnp1m, pdnp1m, pnm, pdnm, pnm1m, pdnm1, ximin);
  LD det = pnp * pnm1m - pnm * pnm1p;
  LD rp = -pnp1p, rm = -pnp1m;
  LD a = (rp * pnm1m - rm * pnm1p) / det;
  LD bb = (rm * pnp - rp * pnm) / det;
  x[0] = ximax;
  x[Np] = ximin;
  int nh = (Np + 1) / 2;
  // gllbasis.ipp:452-456: the reference's literal pi and libm trig seed
  // the Newton iteration; the literal and 2.0*Np_+1.0 are double, so the
  // quotient is formed in double and then held in long double.
  LD dth = 3.1415926535898 / (2.0 * Np + 1.0);
  LD cd = std::cos(2.0L * dth), sd = std::sin(2.0L * dth);
  LD cs = std::cos(dth), ss = std::sin(dth);
  for (int j = 1; j < nh + 1; j++) {
    LD xx = cs;
    for (int k = 0; k < kstop; k++) {
      LD pnp1, pdnp1, pn, pdn, pnm1, pdnm1b;
      jacobi(np1, alpha, beta, pnp1, pdnp1, pn, pdn, pnm1, pdnm1b, xx);
      LD poly = pnp1 + a * pn + bb * pnm1;
      LD pder = pdnp1 + a * pdn + bb * pdnm1b;
      LD recsum = 0.0L;
      for (int i = 0; i < j; i++) recsum += (1.0L / (xx - x[i]));
      LD delx = -poly / (pder - recsum * poly);
      xx += delx;
      LD error = std::fabs(delx) / std::fabs(xx);
      if (error < eps) break;
    }
    x[j] = xx;
    LD cssave = cs * cd - ss * sd;
    ss = cs * sd + ss * cd;
    cs = cssave;
  }
  for (int i = 0; i < nh; i++) x[np1 - i - 1] = -x[i];
  if (Np == (2 * (Np / 2))) x[nh] = 0.0L;
  for (int i = 0; i < nh; i++) {
    LD t = x[Np - i];
    x[Np - i] = x[i];
    x[i] = t;
  }
  // computeWeights (gllbasis.ipp:514-531)
  LD fact = 2.0 / (Np * (Np + 1.0));  // double operands (gllbasis.ipp:520)
  for (int i = 0; i < Np + 1; i++) {
    LD ppn, pder, pm1, pdm1, pm2, pdm2;
    jacobi(Np, alpha, beta, ppn, pder, pm1, pdm1, pm2, pdm2, x[i]);
    Pn[i] = ppn;
    dPn[i] = pder;
    w[i] = (Pn[i] == 0.0L) ? 0.0L : fact / (Pn[i] * Pn[i]);
  }
  // computeDerivMatrix (gllbasis.ipp:579-607)
  std::vector<LD> dphi((size_t)np1 * np1, 0.0L);
  for (int j = 0; j < Np + 1; j++) {
    for (int l = 0; l < j; l++) {
      LD delxi = x[l] - x[j];
      dphi[l + (size_t)j * np1] = Pn[l] / (Pn[j] * delxi + ttiny);
    }
    for (int l = j + 1; l < Np + 1; l++) {
      LD delxi = x[l] - x[j];
      dphi[l + (size_t)j * np1] = Pn[l] / (Pn[j] * delxi + ttiny);
    }
  }
  dphi[0] = -0.25L * Np * (Np + 1.0L);
  dphi[Np + (size_t)Np * np1] = -dphi[0];
  b.p = Np;
  b.N = np1;
  b.xl = x;
  b.Pnl = Pn;
  b.xi.resize(np1);
  b.w.resize(np1);
  b.D.resize((size_t)np1 * np1);
  b.DT.resize((size_t)np1 * np1);
  for (int i = 0; i < np1; i++) {
    b.xi[i] = (double)x[i];
    b.w[i] = (double)w[i];
  }
  for (int i = 0; i < np1; i++)
    for (int j = 0; j < np1; j++) {
      b.D[i + (size_t)j * np1] = (double)dphi[i + (size_t)j * np1];
      b.DT[i + (size_t)j * np1] = (double)dphi[j + (size_t)i * np1];
    }
  return true;
}

// ---------------------------------------------------------------------------
// dmxm: C(nai,nbj) = A(nai,naj) B(nbi,nbj), column-major.  The unrolled
// dmxmN kernels form A(i,1)*B(1,j) + A(i,2)*B(2,j) + ... left to right
// (src/cdg/blas/cff_blas.F:846-943, src/cdg/blas/dmxmp.F:19-897).
// ---------------------------------------------------------------------------
void dmxm(double *C, const double *A, size_t nai, const double *B, size_t nbi, size_t nbj) {
  for (size_t j = 0; j < nbj; j++) {
    for (size_t i = 0; i < nai; i++) {
      double s = A[i] * B[j * nbi];
      for (size_t k = 1; k < nbi; k++) s = s + A[i + k * nai] * B[k + j * nbi];
      C[i + j * nai] = s;
    }
  }
}

int g_threads = 1;
#define OR_PFOR _Pragma("omp parallel for num_threads(g_threads) schedule(static)")
#define OR_PFOR_EL OR_PFOR

}  // namespace

struct OrGrid;
void or_embed_G(OrGrid *g, int i, int j, std::vector<double> &G);

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
struct OrGrid {
  int dim = 3;
  int gtype = 0;  // 0 GE_REGULAR, 1 GE_DEFORMED, 2 GE_2DEMBEDDED (src/cdg/include/gtypes.h:171)
  int N[3] = {1, 1, 1};
  Basis basis[3];
  size_t nelems = 0, ndof = 0, nnodes = 0;  // nnodes per element
  std::vector<double> xn[3];                // node coordinates
  std::vector<double> xp[3];                // periodized coordinates
  std::vector<double> dxidx[3][3];          // dXi_i/dX_j per node
  std::vector<double> jac, mask, mass, imass, imult;
  std::vector<std::vector<size_t>> group;   // GGFX recv_map_[my_rank]: id -> matched set
  std::vector<std::vector<size_t>> buf;     // doOp reduction buffer sources per node
  double minnodedist = 0, volume = 0;
  int nxy() const { return gtype == 2 ? dim + 1 : dim; }
};

namespace {

// GGFX::init local matching (src/cdg/comm/ggfx.hpp:167-245): every node gets
// the box [x - tol, x + tol]; the group of an unmapped id is the sorted set of
// all ids whose boxes intersect its box (Box Intersects,
// src/tbox/spatial/bound/box.hpp:272-276), and that set is assigned to every
// member.  Candidate search here uses a sort on the first coordinate with a
// window wider than the predicate; the predicate itself is the reference's.
void ggfx_match(OrGrid &g, double tol) {
  size_t n = g.ndof;
  int nd = 3;
  std::vector<double> mn[3], mx[3];
  for (int d = 0; d < nd; d++) {
    mn[d].resize(n);
    mx[d].resize(n);
    for (size_t i = 0; i < n; i++) {
      double v = d < g.nxy() ? g.xp[d][i] : 0.0;
      mn[d][i] = v - tol;
      mx[d][i] = v + tol;
    }
  }
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  const std::vector<double> &x0 = g.xp[0];
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return x0[a] < x0[b]; });
  std::vector<double> xs(n);
  for (size_t i = 0; i < n; i++) xs[i] = x0[ord[i]];
  g.group.assign(n, std::vector<size_t>());
  std::vector<char> mapped(n, 0);
  for (size_t id = 0; id < n; id++) {
    if (mapped[id]) continue;
    double lo = x0[id] - 4.0 * tol, hi = x0[id] + 4.0 * tol;
    size_t a = std::lower_bound(xs.begin(), xs.end(), lo) - xs.begin();
    std::set<size_t> ids;
    for (size_t s = a; s < n && xs[s] <= hi; s++) {
      size_t c = ord[s];
      bool hit = true;
      for (int d = 0; d < nd && hit; d++)
        hit = (mn[d][id] <= mx[d][c]) && (mx[d][id] >= mn[d][c]);
      if (hit) ids.insert(c);
    }
    std::vector<size_t> v(ids.begin(), ids.end());
    for (size_t m : v) {
      g.group[m] = v;
      mapped[m] = 1;
    }
  }
}

// GGrid::find_min_dist (src/cdg/grid/ggrid.ipp:977-1013)
double find_min_dist(const OrGrid &g) {
  double tiny = 1000.0 * std::numeric_limits<double>::epsilon();
  double xmin = std::numeric_limits<double>::max();
  int nxy = g.nxy();
  for (size_t j = 0; j + 1 < g.ndof; j++) {
    double s = 0;
    double dd[3] = {0, 0, 0};
    for (int i = 0; i < nxy; i++) dd[i] = g.xn[i][j + 1] - g.xn[i][j];
    s = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
    double xn = std::sqrt(s);
    if (xn > tiny) xmin = std::min(xmin, xn);
  }
  return xmin;
}

void finish_grid(OrGrid &g, double tol_fac);

}  // namespace

extern "C" {

int or_gll_basis(int p, double *xi, double *w, double *D) {
  Basis b;
  if (!gll(p, b)) return 1;
  std::copy(b.xi.begin(), b.xi.end(), xi);
  std::copy(b.w.begin(), b.w.end(), w);
  std::copy(b.D.begin(), b.D.end(), D);
  return 0;
}

void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

// Element-local and pointwise passes run on g_threads OpenMP threads over
// contiguous index ranges (the reference's element decomposition over ranks
// splits the same work); every output entry is computed by the same
// operations as with one thread, so results do not depend on g_threads.

void or_dmxm(double *C, const double *A, size_t nai, size_t naj, const double *B, size_t nbi,
             size_t nbj) {
  (void)naj;  // dmxm ignores naj (src/cdg/blas/cff_blas.F:880-927)
  dmxm(C, A, nai, B, nbi, nbj);
}

// GGridBox element generation: find_rank_subdomain (single rank) and
// do_elems2d/do_elems3d (src/cdg/grid/ggrid_box.ipp:862-978, 254-463);
// GElem_base::dogeom2d/3d for GE_REGULAR (src/cdg/sem/gelem_base.cpp:877-1048);
// boundary detection by coordinate (find_gbdy_ind2d, ggrid_box.ipp:1494-1600)
// and periodize (ggrid_box.ipp:753-802).
OrGrid *or_grid_box_create(int dim, const int *p, const int *ne, const double *P0,
                           const double *P1, const int *bdy, double tol_fac) {
  if (dim != 2 && dim != 3) return nullptr;
  OrGrid *gp = new OrGrid();
  OrGrid &g = *gp;
  g.dim = dim;
  g.gtype = 0;
  for (int d = 0; d < 3; d++) g.N[d] = 1;
  for (int d = 0; d < dim; d++) {
    if (!gll(p[d], g.basis[d])) {
      delete gp;
      return nullptr;
    }
    g.N[d] = p[d] + 1;
  }
  g.nnodes = (size_t)g.N[0] * g.N[1] * g.N[2];
  g.nelems = 1;
  for (int d = 0; d < dim; d++) g.nelems *= (size_t)ne[d];
  g.ndof = g.nelems * g.nnodes;
  double dx[3] = {0, 0, 0}, Lbox[3];
  for (int d = 0; d < dim; d++) {
    Lbox[d] = P1[d] - P0[d];
    dx[d] = Lbox[d] / (double)ne[d];
  }
  for (int d = 0; d < 3; d++) {
    g.xn[d].assign(d < dim ? g.ndof : 0, 0.0);
    for (int j = 0; j < 3; j++) g.dxidx[d][j].assign(g.ndof, 0.0);
  }
  g.jac.assign(g.ndof, 0.0);
  size_t n = 0;
  // Vertex sets in the order of hmesh_/qmesh_ (ggrid_box.ipp:887-975).
  static const int off2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  static const int off3[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  int nk = dim == 3 ? ne[2] : 1;
  for (long k = 0; k < nk; k++)
    for (long j = 0; j < ne[1]; j++)
      for (long i = 0; i < ne[0]; i++) {
        double v0[3] = {P0[0] + i * dx[0], P0[1] + j * dx[1], dim == 3 ? P0[2] + k * dx[2] : 0.0};
        int nv = dim == 3 ? 8 : 4;
        double V[8][3];
        for (int m = 0; m < nv; m++)
          for (int l = 0; l < dim; l++) {
            int o = dim == 3 ? off3[m][l] : off2[m][l];
            double dv = o ? dx[l] : 0.0;
            V[m][l] = v0[l] + dv;
          }
        size_t ebeg = n;
        // Linear shape function node coordinates (gshapefcn_linear.ipp:171-348):
        // x_l = sum_m N_m(xi) * (v_m[l] * 2^-dim), accumulated from zero.
        double fac = dim == 3 ? 0.125 : 0.25;
        const std::vector<double> &x0 = g.basis[0].xi, &x1 = g.basis[1].xi;
        for (int l = 0; l < dim; l++) {
          size_t nn = ebeg;
          for (int kk = 0; kk < g.N[2]; kk++)
            for (int jj = 0; jj < g.N[1]; jj++)
              for (int ii = 0; ii < g.N[0]; ii++, nn++) {
                double acc = 0.0;
                for (int m = 0; m < nv; m++) {
                  double rm = 1.0 - x0[ii], rp = 1.0 + x0[ii];
                  double sm = 1.0 - x1[jj], sp = 1.0 + x1[jj];
                  double Ni;
                  if (dim == 2) {
                    switch (m) {
                      case 0: Ni = rm * sm; break;
                      case 1: Ni = rp * sm; break;
                      case 2: Ni = rp * sp; break;
                      default: Ni = rm * sp; break;
                    }
                  } else {
                    double tm = 1.0 - g.basis[2].xi[kk], tp = 1.0 + g.basis[2].xi[kk];
                    switch (m) {
                      case 0: Ni = rm * sm * tm; break;
                      case 1: Ni = rp * sm * tm; break;
                      case 2: Ni = rp * sp * tm; break;
                      case 3: Ni = rm * sp * tm; break;
                      case 4: Ni = rm * sm * tp; break;
                      case 5: Ni = rp * sm * tp; break;
                      case 6: Ni = rp * sp * tp; break;
                      default: Ni = rm * sp * tp; break;
                    }
                  }
                  acc += Ni * (V[m][l] * fac);
                }
                g.xn[l][nn] = acc;
              }
        }
        // dogeom for GE_REGULAR: L from vertices, dXi/dX = 2/L, Jac = prod(L)/2^d
        double L[3];
        L[0] = V[1][0] - V[0][0];
        L[1] = V[2][1] - V[1][1];
        if (dim == 3) L[2] = V[4][2] - V[0][2];
        double jv = dim == 3 ? 0.125 * L[0] * L[1] * L[2] : 0.25 * L[0] * L[1];
        for (size_t q = 0; q < g.nnodes; q++) {
          for (int d = 0; d < dim; d++) g.dxidx[d][0][ebeg + q] = 2.0 / L[d];
          g.jac[ebeg + q] = jv;
        }
        n += g.nnodes;
      }
  g.minnodedist = find_min_dist(g);
  // Boundary bins and periodize (ggrid_box.ipp:753-802): a node that lies on
  // any periodic face has every coordinate that sits on P1 moved to P0.
  double eps = 0.125 * g.minnodedist;
  auto fuzzyeq = [](double x, double y, double e) { return (y <= (x + e)) && (y >= (x - e)); };
  g.mask.assign(g.ndof, 1.0);
  for (int d = 0; d < 3; d++) g.xp[d] = g.xn[d];
  for (size_t q = 0; q < g.ndof; q++) {
    bool onper = false, ondir = false;
    for (int d = 0; d < dim; d++) {
      for (int s = 0; s < 2; s++) {
        double P = s ? P1[d] : P0[d];
        if (fuzzyeq(P, g.xn[d][q], eps)) {
          int bt = bdy[2 * d + s];
          if (bt == 6) onper = true;
          else if (bt != 7) ondir = true;
        }
      }
    }
    // init_bc_info (src/cdg/grid/ggrid.ipp:1459-1470)
    if (ondir) g.mask[q] = 0.0;
    if (onper) {
      for (int d = 0; d < dim; d++)
        if (fuzzyeq(P1[d], g.xn[d][q], eps)) g.xp[d][q] = P0[d];
    }
  }
  finish_grid(g, tol_fac);
  return gp;
}

void or_grid_destroy(OrGrid *g) { delete g; }
size_t or_grid_ndof(const OrGrid *g) { return g->ndof; }
size_t or_grid_nelems(const OrGrid *g) { return g->nelems; }
int or_grid_dim(const OrGrid *g) { return g->dim; }
int or_grid_gtype(const OrGrid *g) { return g->gtype; }
double or_grid_minnodedist(const OrGrid *g) { return g->minnodedist; }
double or_grid_volume(const OrGrid *g) { return g->volume; }

int or_grid_get(const OrGrid *g, const char *which, double *out) {
  std::string w(which);
  const std::vector<double> *src = nullptr;
  if (w == "x") src = &g->xn[0];
  else if (w == "y") src = &g->xn[1];
  else if (w == "z") src = &g->xn[2];
  else if (w == "xp") src = &g->xp[0];
  else if (w == "yp") src = &g->xp[1];
  else if (w == "zp") src = &g->xp[2];
  else if (w == "mask") src = &g->mask;
  else if (w == "mass") src = &g->mass;
  else if (w == "imass") src = &g->imass;
  else if (w == "jac") src = &g->jac;
  else if (w == "imult") src = &g->imult;
  else if (w.size() == 8 && w.compare(0, 6, "dxidx_") == 0) {
    int i = w[6] - '0', j = w[7] - '0';
    if (i < 0 || i > 2 || j < 0 || j > 2) return 1;
    src = &g->dxidx[i][j];
  }
  if (!src || src->size() != g->ndof) return 1;
  std::copy(src->begin(), src->end(), out);
  return 0;
}

int or_grid_ggfx(const OrGrid *g, size_t *offs, size_t *members, size_t *nmem) {
  size_t tot = 0;
  for (size_t i = 0; i < g->ndof; i++) tot += g->group[i].size();
  if (nmem) *nmem = tot;
  if (!members) return 0;
  size_t c = 0;
  for (size_t i = 0; i < g->ndof; i++) {
    offs[i] = c;
    for (size_t m : g->group[i]) members[c++] = m;
  }
  offs[g->ndof] = c;
  return 0;
}

// ---------------------------------------------------------------------------
// Reference-space derivatives (GGrid::grefderiv_constp, ggrid.ipp:2213-2265
// and the GMTK batched forms, gmtk.ipp:1404-1844, 1998-2294).  Directions 2
// and 3 receive getDerivMatrix(!dotrans) and contract as U * D^T, so every
// direction computes y(out) = sum_k Dop(out,k) u(k) with k ascending.
// ---------------------------------------------------------------------------
void or_grefderiv_raw(int dim, const int *N, size_t ne, const double *const *Dm, int idir,
                      int dotrans, const double *u, double *du) {
  int N0 = N[0], N1 = N[1], N2 = dim == 3 ? N[2] : 1;
  size_t nn = (size_t)N0 * N1 * N2;
  const double *D = Dm[idir - 1];
  int Nd = N[idir - 1];
  // op(o,k): derivative matrix entry used for output index o, input index k.
  std::vector<double> op((size_t)Nd * Nd);
  for (int o = 0; o < Nd; o++)
    for (int k = 0; k < Nd; k++) op[o + (size_t)k * Nd] = dotrans ? D[k + (size_t)o * Nd] : D[o + (size_t)k * Nd];
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long e = 0; e < (long)ne; e++) {
    const double *ue = u + e * nn;
    double *de = du + e * nn;
    if (idir == 1) {
      // I x I x D1: dmxm(y, D1, N0, u, N0, N1*N2)
      dmxm(de, op.data(), N0, ue, N0, (size_t)N1 * N2);
    } else if (idir == 2) {
      // I x D2 x I: per z-slice dmxm(y, u, N0, D2T, N1, N1) with D2T(k,j)=op(j,k)
      std::vector<double> BT((size_t)N1 * N1);
      for (int k = 0; k < N1; k++)
        for (int j = 0; j < N1; j++) BT[k + (size_t)j * N1] = op[j + (size_t)k * N1];
      for (int s = 0; s < N2; s++)
        dmxm(de + (size_t)s * N0 * N1, ue + (size_t)s * N0 * N1, N0, BT.data(), N1, N1);
    } else {
      std::vector<double> BT((size_t)N2 * N2);
      for (int k = 0; k < N2; k++)
        for (int j = 0; j < N2; j++) BT[k + (size_t)j * N2] = op[j + (size_t)k * N2];
      dmxm(de, ue, (size_t)N0 * N1, BT.data(), N2, N2);
    }
  }
}

void or_grefderiv(OrGrid *g, const double *u, int idir, int dotrans, double *du) {
  const double *Dm[3] = {g->basis[0].D.data(), g->basis[1].D.data(), g->basis[2].D.data()};
  if (g->dim == 2 && idir == 3) {
    // grefderiv_varp 2-D idir 3 yields zero (ggrid.ipp:2133-2137)
    std::fill(du, du + g->ndof, 0.0);
    return;
  }
  or_grefderiv_raw(g->dim, g->N, g->nelems, Dm, idir, dotrans, u, du);
}

// compute_grefdiv (ggrid.ipp:1847-1927): divu = 0; divu += D1 u1; += D2 u2;
// += D3 u3.  The reference takes the 3-D direction-3 matrix from basis 1
// (ggrid.ipp:1906); the contraction is only well formed when N2 == N3, in
// which case both bases are the same GLL basis, so basis 2 is used here.
void or_grefdiv(OrGrid *g, const double *const *u, int dotrans, double *divu) {
  size_t n = g->ndof;
  std::vector<double> et(n);
  std::fill(divu, divu + n, 0.0);
  for (int k = 0; k < g->dim; k++) {
    if (!u[k]) continue;
    or_grefderiv(g, u[k], k + 1, dotrans, et.data());
    OR_PFOR
    for (size_t j = 0; j < n; j++) divu[j] += et[j];
  }
}

// GGrid::deriv (ggrid.ipp:1163-1191)
void or_deriv(OrGrid *g, const double *u, int idir, double *du) {
  size_t n = g->ndof;
  if (g->gtype == 0) {
    or_grefderiv(g, u, idir, 0, du);
    const std::vector<double> &f = g->dxidx[idir - 1][0];
    OR_PFOR
    for (size_t j = 0; j < n; j++) du[j] *= f[j];
  } else {
    std::vector<double> tmp(n);
    or_grefderiv(g, u, 1, 0, du);
    const std::vector<double> &f0 = g->dxidx[0][idir - 1];
    OR_PFOR
    for (size_t j = 0; j < n; j++) du[j] *= f0[j];
    for (int jd = 1; jd < g->dim; jd++) {
      or_grefderiv(g, u, jd + 1, 0, tmp.data());
      const std::vector<double> &f = g->dxidx[jd][idir - 1];
      OR_PFOR
      for (size_t j = 0; j < n; j++) tmp[j] *= f[j];
      OR_PFOR
      for (size_t j = 0; j < n; j++) du[j] += tmp[j];
    }
  }
}

// GGrid::integrate (ggrid.ipp:1077-1148): tmp = W2 W1 W0 u (left to right),
// tmp *= Jac, serial sum.
double or_integrate(OrGrid *g, const double *u) {
  size_t n = 0;
  double s = 0.0;
  const std::vector<double> &W0 = g->basis[0].w, &W1 = g->basis[1].w;
  for (size_t e = 0; e < g->nelems; e++) {
    if (g->dim == 2) {
      for (int k = 0; k < g->N[1]; k++)
        for (int j = 0; j < g->N[0]; j++, n++) {
          double t = W1[k] * W0[j] * u[n];
          t = t * g->jac[n];
          s += t;
        }
    } else {
      const std::vector<double> &W2 = g->basis[2].w;
      for (int k = 0; k < g->N[2]; k++)
        for (int j = 0; j < g->N[1]; j++)
          for (int i = 0; i < g->N[0]; i++, n++) {
            double t = W2[k] * W1[j] * W0[i] * u[n];
            t = t * g->jac[n];
            s += t;
          }
    }
  }
  return s;
}

void or_mass_apply(OrGrid *g, const double *in, double *out) {
  // GMass::opVec_prod: mass_.pointProd(input, output) (gmass.ipp:225-232)
  OR_PFOR
  for (size_t j = 0; j < g->ndof; j++) out[j] = g->mass[j] * in[j];
}

void or_imass_apply(OrGrid *g, const double *in, double *out) {
  OR_PFOR
  for (size_t j = 0; j < g->ndof; j++) out[j] = g->imass[j] * in[j];
}

// GGFX::doOp (ggfx.hpp:344-443) on a single rank: for every local_id in
// ascending order (recv_map_ is a std::map) and every id in its matched set,
// u[local_id] is appended to id's buffer (ggfx.hpp:401-405).  Sum and Smooth
// accumulate the buffer from zero (ggfx.hpp:60-76), Max/Min from
// lowest()/max().
void or_dss(OrGrid *g, int op, double *u) {
  size_t n = g->ndof;
  std::vector<double> out(n);
  OR_PFOR
  for (size_t i = 0; i < n; i++) {
    const std::vector<size_t> &grp = g->buf[i];
    double r;
    if (op == 0 || op == 1) {
      r = 0.0;
      for (size_t m : grp) r = r + u[m];
      if (op == 1) r = r / (double)grp.size();
    } else if (op == 2) {
      r = std::numeric_limits<double>::lowest();
      for (size_t m : grp) r = std::max(r, u[m]);
    } else {
      r = std::numeric_limits<double>::max();
      for (size_t m : grp) r = std::min(r, u[m]);
    }
    out[i] = r;
  }
  std::copy(out.begin(), out.end(), u);
}

// GGFX::get_mult/get_imult (ggfx.hpp:451-487): count of buffer sources.
void or_imult(OrGrid *g, double *imult) {
  for (size_t i = 0; i < g->ndof; i++) imult[i] = 1.0 / (double)g->buf[i].size();
}

// GHelmholtz reg_prod / def_prod / embed_prod (ghelmholtz.ipp:76-378)
void or_helmholtz_apply(OrGrid *g, const double *p, size_t np, const double *q, const double *u,
                        double *uo) {
  size_t n = g->ndof;
  int dim = g->dim;
  std::vector<std::vector<double>> gdu(dim, std::vector<double>(n));
  std::vector<double> tmp(n);
  bool pfull = p && np >= n;
  if (g->gtype == 0) {
    for (int k = 0; k < dim; k++) or_grefderiv(g, u, k + 1, 0, gdu[k].data());
    for (int k = 0; k < dim; k++) {
      const std::vector<double> &dx = g->dxidx[k][0];
      OR_PFOR
      for (size_t j = 0; j < n; j++) {
        double G = dx[j] * dx[j];  // reg_init (ghelmholtz.ipp:571-575)
        gdu[k][j] *= G;
      }
      if (pfull) {
        OR_PFOR
        for (size_t j = 0; j < n; j++) gdu[k][j] *= p[j];
      }
      OR_PFOR
      for (size_t j = 0; j < n; j++) gdu[k][j] = g->mass[j] * gdu[k][j];
    }
  } else {
    // def_prod/embed_prod: gdu_j = sum_i G(i,j) D_i u, with G from def_init
    // (ghelmholtz.ipp:415-519); G is built in finish_grid for embedded grids.
    std::vector<std::vector<double>> du(dim, std::vector<double>(n));
    for (int k = 0; k < dim; k++) or_grefderiv(g, u, k + 1, 0, du[k].data());
    std::vector<double> G;
    for (int jj = 0; jj < dim; jj++) {
      std::fill(gdu[jj].begin(), gdu[jj].end(), 0.0);
      for (int ii = 0; ii < dim; ii++) {
        or_embed_G(g, ii, jj, G);
        OR_PFOR
        for (size_t j = 0; j < n; j++) gdu[jj][j] += du[ii][j] * G[j];
      }
    }
    if (pfull)
      for (int k = 0; k < dim; k++) {
        OR_PFOR
        for (size_t j = 0; j < n; j++) gdu[k][j] *= p[j];
      }
  }
  const double *up[3] = {gdu[0].data(), gdu[1].data(), dim == 3 ? gdu[2].data() : nullptr};
  or_grefdiv(g, up, 1, uo);
  if (p && np < n && p[0] != 1.0) {
    OR_PFOR
    for (size_t j = 0; j < n; j++) uo[j] *= p[0];
  }
  if (q) {
    OR_PFOR
    for (size_t j = 0; j < n; j++) {
      double t = g->mass[j] * u[j];
      t *= q[j];
      uo[j] += t;
    }
  }
}

// Assembled diagonal of (p L + q M) for the Jacobi preconditioner (a
// build-side addition; the reference has no Jacobi preconditioner), from the
// closed form of the element operator sum_{a,b} D_a^T diag(G_ab) D_b + q M,
// D_a being the derivative along direction a (D(l, n_a) on the line of
// nodes through n).  Its diagonal entry at node n is
//   sum_a sum_l D(l, n_a)^2 G_aa(line_a(n, l))
//     + sum_{a != b} D(n_a, n_a) D(n_b, n_b) G_ab(n)      (times p)  + q M(n).
// On regular elements G_ab = 0 for a != b and G_aa = M (dXi_a/dx_a)^2, so the
// line sum becomes G_aa Jac prod_{o != a} W(n_o) sum_l D(l, n_a)^2 W(l).  On
// meshes with node metrics G_ab is the vector of or_embed_G (3-D: with the
// doubled off-diagonals the operator applies).  The element diagonals are
// assembled with a GGFX Sum.
void or_helmholtz_diag_inv(OrGrid *g, const double *p, size_t np, const double *q, double *dinv) {
  size_t n = g->ndof;
  std::vector<double> d(n, 0.0);
  int N0 = g->N[0], N1 = g->N[1], N2 = g->N[2];
  size_t nn = g->nnodes;
  const int dim = g->dim;
  const bool pfull = p && np >= n;
  std::vector<std::vector<double>> G(9);
  if (g->gtype != 0)
    for (int b = 0; b < dim; b++)
      for (int a = 0; a <= b; a++) or_embed_G(g, a, b, G[a * 3 + b]);
  const size_t stride[3] = {1, (size_t)N0, (size_t)N0 * N1};
  for (size_t e = 0; e < g->nelems; e++)
    for (int k = 0; k < N2; k++)
      for (int j = 0; j < N1; j++)
        for (int i = 0; i < N0; i++) {
          size_t id = e * nn + (size_t)k * N0 * N1 + (size_t)j * N0 + i;
          int idx[3] = {i, j, k};
          double lap = 0.0;
          if (g->gtype == 0) {
            double Wn[3] = {g->basis[0].w[i], g->basis[1].w[j], dim == 3 ? g->basis[2].w[k] : 1.0};
            for (int dd = 0; dd < dim; dd++) {
              const Basis &b = g->basis[dd];
              int Nd = b.N;
              // p (a full field) multiplies pointwise before D^T, so it
              // enters the line sum at each line node m
              size_t line = id - (size_t)idx[dd] * stride[dd];
              double s = 0.0;
              for (int m = 0; m < Nd; m++) {
                double Dm = b.D[m + (size_t)idx[dd] * Nd];
                double wm = pfull ? b.w[m] * p[line + (size_t)m * stride[dd]] : b.w[m];
                s += Dm * Dm * wm;
              }
              double other = 1.0;
              for (int o = 0; o < dim; o++)
                if (o != dd) other *= Wn[o];
              double Gd = g->dxidx[dd][0][id] * g->dxidx[dd][0][id];
              lap += Gd * s * other * g->jac[id];
            }
          } else {
            for (int a = 0; a < dim; a++) {
              const Basis &b = g->basis[a];
              const std::vector<double> &Ga = G[a * 3 + a];
              size_t line = id - (size_t)idx[a] * stride[a];
              double s = 0.0;
              for (int l = 0; l < b.N; l++) {
                double Dl = b.D[l + (size_t)idx[a] * b.N];
                size_t ql = line + (size_t)l * stride[a];
                double gv = pfull ? p[ql] * Ga[ql] : Ga[ql];
                s += Dl * Dl * gv;
              }
              lap += s;
            }
            for (int a = 0; a < dim; a++)
              for (int b = a + 1; b < dim; b++) {
                double Daa = g->basis[a].D[idx[a] * (g->basis[a].N + 1)];
                double Dbb = g->basis[b].D[idx[b] * (g->basis[b].N + 1)];
                double gv = pfull ? p[id] * G[a * 3 + b][id] : G[a * 3 + b][id];
                lap += 2.0 * Daa * Dbb * gv;
              }
          }
          double pv = p ? (pfull ? 1.0 : p[0]) : 1.0;
          double val = pv * lap;
          if (q) val += q[id] * g->mass[id];
          d[id] = val;
        }
  or_dss(g, 0, d.data());
  for (size_t j = 0; j < n; j++) dinv[j] = 1.0 / d[j];
}

// GLLBasis::evalBasis (3), the matrix form (gllbasis.ipp:1341-1377):
// phi_j(eta) without the ttiny shifts of the vector form.
double evalBasisM(const Basis &b, int j, double eta) {
  LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  int Np = b.p;
  LD fact = -1.0 / (Np * (Np + 1.0)), xi = static_cast<LD>(eta);
  double fRet = 1.0;
  if (xi < -1.0L || xi > 1.0L) fRet = 0.0;
  else if (std::fabs(xi - b.xl[j]) > tetiny) {
    LD ppn_j = b.Pnl[j], ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2;
    jacobi(Np, 0.0L, 0.0L, ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2, xi);
    LD gfact = (1.0L - xi * xi) / (xi - b.xl[j]);
    fRet = static_cast<double>(fact * gfact * pder_xi / ppn_j);
  }
  return fRet;
}

// GGrid::dealias (ggrid.ipp:2424-2503) with init_qdealias (2351-2414), for
// regular isotropic grids: v1 and v2 interpolated to the order-pq GLL grid of
// each element by IQP(a, i) = phi_i(xq_a) applied direction by direction
// (x, then y, then z, each a dmxm-order contraction), multiplied pointwise
// and by the q tensor weights, projected back with IQP^T in the same
// direction order, and multiplied by iWp = Jac * imass.  The contractions
// follow GMTK::D2_X_D1 / D3_X_D2_X_D1 (gmtk.ipp:1570-1600, 1930-1992) as those
// routines intend for non-square factors: in 3-D the middle (y) pass runs
// over every z plane of its input with that input's plane stride.
void or_dealias(OrGrid *g, int pq, const double *v1, const double *v2, double *prod) {
  const int dim = g->dim, N = g->N[0], Q = pq + 1;
  Basis bq;
  gll(pq, bq);
  std::vector<double> I((size_t)Q * N);  // I[a + Q*i] = phi_i(xq_a)
  for (int a = 0; a < Q; a++)
    for (int i = 0; i < N; i++) I[a + (size_t)Q * i] = evalBasisM(g->basis[0], i, bq.xi[a]);
  const int NQd = dim == 3 ? Q * Q * Q : Q * Q, NNd = dim == 3 ? N * N * N : N * N;
  // contraction along one direction: in is (n0 x n1 x n2) x-fastest, the
  // direction d has length lin in and lout out, op(o, k) the matrix entry
  auto pass = [](const double *in, double *out, int n0, int n1, int n2, int d, int lout,
                 const std::function<double(int, int)> &op) {
    const int lin = d == 0 ? n0 : (d == 1 ? n1 : n2);
    const int m0 = d == 0 ? lout : n0, m1 = d == 1 ? lout : n1, m2 = d == 2 ? lout : n2;
    for (int c = 0; c < m2; c++)
      for (int b = 0; b < m1; b++)
        for (int a = 0; a < m0; a++) {
          double acc = 0.0;
          for (int k = 0; k < lin; k++) {
            const int ia = d == 0 ? k : a, ib = d == 1 ? k : b, ic = d == 2 ? k : c;
            acc += op(d == 0 ? a : (d == 1 ? b : c), k) * in[ia + n0 * (ib + n1 * ic)];
          }
          out[a + m0 * (b + m1 * c)] = acc;
        }
  };
  auto fwd = [&](int o, int k) { return I[o + (size_t)Q * k]; };  // IQP(o, k)
  auto bwd = [&](int o, int k) { return I[k + (size_t)Q * o]; };  // IQP^T(o, k)
  const int z = dim == 3 ? 1 : 0;
  OR_PFOR_EL
  for (long e = 0; e < (long)g->nelems; e++) {
    std::vector<double> a1((size_t)NQd), a2((size_t)NQd), b1((size_t)NQd), b2((size_t)NQd);
    for (int f = 0; f < 2; f++) {
      const double *v = (f == 0 ? v1 : v2) + (size_t)e * NNd;
      std::vector<double> &dst = f == 0 ? b1 : b2;
      pass(v, a1.data(), N, N, z ? N : 1, 0, Q, fwd);
      if (z) {
        pass(a1.data(), a2.data(), Q, N, N, 1, Q, fwd);
        pass(a2.data(), dst.data(), Q, Q, N, 2, Q, fwd);
      } else {
        pass(a1.data(), dst.data(), Q, N, 1, 1, Q, fwd);
      }
    }
    // pointwise product times the q tensor weights ((Wx Wy) Wz)
    for (int c = 0; c < (z ? Q : 1); c++)
      for (int b = 0; b < Q; b++)
        for (int a = 0; a < Q; a++) {
          const size_t n = a + (size_t)Q * (b + (size_t)Q * c);
          double w = bq.w[a] * bq.w[b];
          if (z) w = w * bq.w[c];
          b1[n] = b1[n] * b2[n];
          b1[n] = b1[n] * w;
        }
    double *out = prod + (size_t)e * NNd;
    pass(b1.data(), a1.data(), Q, Q, z ? Q : 1, 0, N, bwd);
    if (z) {
      pass(a1.data(), a2.data(), N, Q, Q, 1, N, bwd);
      pass(a2.data(), out, N, N, Q, 2, N, bwd);
    } else {
      pass(a1.data(), out, N, Q, 1, 1, N, bwd);
    }
    for (int q = 0; q < NNd; q++) {
      const size_t gi = (size_t)e * NNd + q;
      out[q] *= g->jac[gi] * g->imass[gi];
    }
  }
}

// GAdvect::apply (gadvect.ipp:57-99) for ivec = -2 (no Neumann hack).
void or_advect(OrGrid *g, const double *p, int nu, const double *const *u, double *po) {
  size_t n = g->ndof;
  std::vector<double> t0(n);
  if (u[0]) {
    or_deriv(g, p, 1, po);
    OR_PFOR
    for (size_t j = 0; j < n; j++) po[j] *= u[0][j];
  } else {
    std::fill(po, po + n, 0.0);
  }
  for (int jv = 1; jv < nu; jv++) {
    if (!u[jv]) continue;
    or_deriv(g, p, jv + 1, t0.data());
    OR_PFOR
    for (size_t j = 0; j < n; j++) t0[j] *= u[jv][j];
    OR_PFOR
    for (size_t j = 0; j < n; j++) po[j] += t0[j];
  }
  OR_PFOR
  for (size_t j = 0; j < n; j++) po[j] *= g->mass[j];
}

// GBurgers::dudt_impl (gburgers.ipp:202-257)
void or_burgers_rhs(OrGrid *g, int mode, double nu, int nstate, const double *const *u,
                    double *const *dudt) {
  size_t n = g->ndof;
  std::vector<double> urhs(n);
  double pnu = nu;
  if (mode == 2) {
    for (int k = 0; k < nstate; k++) {
      or_helmholtz_apply(g, &pnu, 1, nullptr, u[k], urhs.data());
      OR_PFOR
      for (size_t j = 0; j < n; j++) urhs[j] *= -1.0;
      or_imass_apply(g, urhs.data(), dudt[k]);
    }
    return;
  }
  if (mode == 1) {
    or_advect(g, u[0], nstate - 1, u + 1, dudt[0]);
    or_helmholtz_apply(g, &pnu, 1, nullptr, u[0], urhs.data());
    OR_PFOR
    for (size_t j = 0; j < n; j++) urhs[j] = -1.0 * urhs[j] + -1.0 * dudt[0][j];
    or_imass_apply(g, urhs.data(), dudt[0]);
    return;
  }
  for (int k = 0; k < g->dim; k++) {
    or_advect(g, u[k], nstate, u, dudt[k]);
    or_helmholtz_apply(g, &pnu, 1, nullptr, u[k], urhs.data());
    OR_PFOR
    for (size_t j = 0; j < n; j++) urhs[j] = -1.0 * urhs[j] + -1.0 * dudt[k][j];
    or_imass_apply(g, urhs.data(), dudt[k]);
  }
  // the loop runs over GDIM (gburgers.ipp:247): a sphere grid's third
  // velocity component has no RHS; it is returned as zero
  for (int k = g->dim; k < nstate; k++) std::fill(dudt[k], dudt[k] + n, 0.0);
}

}  // extern "C"

namespace {
// GMTK::constrain2sphere (2) (gmtk.ipp:2956-2980), in place with aliasing.
void constrain2sphere(OrGrid *g, std::vector<double *> &v) {
  if (g->gtype != 2 || v.size() < 3) return;
  size_t n = g->ndof;
  double x = g->xn[0][0], y = g->xn[1][0], z = g->xn[2][0];
  double r2 = x * x + y * y + z * z;
  double ri = 1.0 / r2;
  for (size_t j = 0; j < n; j++) {
    x = g->xn[0][j];
    y = g->xn[1][j];
    z = g->xn[2][j];
    v[0][j] = (v[0][j] * (r2 - x * x) - v[1][j] * x * y - v[2][j] * x * z) * ri;
    v[1][j] = (-v[0][j] * y * x + v[1][j] * (r2 - y * y) - v[2][j] * y * z) * ri;
    v[2][j] = (-v[0][j] * z * x - v[1][j] * z * y + v[2][j] * (r2 - z * z)) * ri;
  }
}

}  // namespace

extern "C" void or_constrain2sphere(OrGrid *g, double *v0, double *v1, double *v2) {
  std::vector<double *> v = {v0, v1, v2};
  constrain2sphere(g, v);
}

namespace {

void butcher(int order, std::vector<double> &alpha, std::vector<double> &beta,
             std::vector<double> &c) {
  // GButcherRK::computeCoeffs (gbutcherrk.ipp:80-127); beta(i,j) at [i*order+j]
  alpha.assign(order, 0.0);
  c.assign(order, 0.0);
  beta.assign((size_t)order * order, 0.0);
  if (order == 1) {
    alpha[0] = 0;
    c[0] = 1.0;
  } else if (order == 2) {
    alpha[0] = 0; c[0] = 0.5;
    alpha[1] = 1.0; c[1] = 0.5;
    beta[1 * order + 0] = 1.0;
  } else if (order == 3) {
    alpha[0] = 0; c[0] = 0.25;
    alpha[1] = 1.0 / 3.0; c[1] = 0.0;
    alpha[2] = 2.0 / 3.0; c[2] = 0.75;
    beta[1 * order + 0] = 1.0 / 3.0;
    beta[2 * order + 1] = 2.0 / 3.0;
  } else {
    alpha[0] = 0; c[0] = 1.0 / 6.0;
    alpha[1] = 0.5; c[1] = 1.0 / 3.0;
    alpha[2] = 0.5; c[2] = 1.0 / 3.0;
    alpha[3] = 1.0; c[3] = 1.0 / 6.0;
    beta[1 * order + 0] = 0.5;
    beta[2 * order + 1] = 0.5;
    beta[3 * order + 2] = 1.0;
  }
}
}  // namespace

extern "C" {

// GBurgers::step_impl -> step_exrk -> GExRKStepper::step_b
// (gburgers.ipp:274-333, 392-407; gexrk_stepper.ipp:251-341).
void or_burgers_step(OrGrid *g, int mode, double nu, int order, int nstate, double t, double dt,
                     double *const *ust) {
  size_t n = g->ndof;
  int nev = mode == 0 ? nstate : 1;  // evolved components
  std::vector<double> alpha, beta, c;
  butcher(order, alpha, beta, c);
  int nstage = order;
  double h = dt;
  std::vector<std::vector<double>> uin(nev, std::vector<double>(n)), u(nev, std::vector<double>(n));
  std::vector<std::vector<std::vector<double>>> K(std::max(nstage - 1, 1),
                                                  std::vector<std::vector<double>>(nev, std::vector<double>(n)));
  std::vector<double> isum(n);
  // uold = uevolve (deep copy); uout = uin
  for (int k = 0; k < nev; k++) std::copy(ust[k], ust[k] + n, uin[k].begin());
  double *const *uout = ust;
  // RHS state: evolved comps (+ prescribed velocities for pure advection)
  auto rhs = [&](std::vector<std::vector<double>> &us, std::vector<std::vector<double>> &Kout) {
    std::vector<const double *> up(nstate);
    std::vector<double *> dp(nev);
    for (int k = 0; k < nev; k++) up[k] = us[k].data();
    for (int k = nev; k < nstate; k++) up[k] = ust[k];
    for (int k = 0; k < nev; k++) dp[k] = Kout[k].data();
    // heat mode evolves only uin[0] (gburgers.ipp:295-300): its RHS sees the
    // evolved component alone
    or_burgers_rhs(g, mode, nu, mode == 2 ? nev : nstate, up.data(), dp.data());
  };
  auto ptrs = [&](std::vector<std::vector<double>> &vv) {
    std::vector<double *> p(vv.size());
    for (size_t k = 0; k < vv.size(); k++) p[k] = vv[k].data();
    return p;
  };
  for (int m = 0; m < nstage - 1; m++) {
    for (int k = 0; k < nev; k++) {
      std::fill(isum.begin(), isum.end(), 0.0);
      for (int j = 0; j < m; j++) {
        double b = beta[(size_t)m * order + j] * h;
        OR_PFOR
        for (size_t q = 0; q < n; q++) isum[q] = 1.0 * isum[q] + b * K[j][k][q];
      }
      OR_PFOR
      for (size_t q = 0; q < n; q++) u[k][q] = uin[k][q] + isum[q];
    }
    {
      std::vector<double *> up = ptrs(u);
      constrain2sphere(g, up);
    }
    rhs(u, K[m]);
    for (int k = 0; k < nev; k++) {
      or_dss(g, 1, K[m][k].data());
      double cm = c[m] * h;
      OR_PFOR
      for (size_t q = 0; q < n; q++) isum[q] = K[m][k][q] * cm;
      OR_PFOR
      for (size_t q = 0; q < n; q++) uout[k][q] += isum[q];
    }
    {
      std::vector<double *> up(uout, uout + nev);
      constrain2sphere(g, up);
    }
  }
  // Final stage
  for (int k = 0; k < nev; k++) {
    std::fill(isum.begin(), isum.end(), 0.0);
    for (int j = 0; j < nstage - 1; j++) {
      double b = beta[(size_t)(nstage - 1) * order + j] * h;
      OR_PFOR
      for (size_t q = 0; q < n; q++) isum[q] = 1.0 * isum[q] + b * K[j][k][q];
    }
    OR_PFOR
    for (size_t q = 0; q < n; q++) u[k][q] = uin[k][q] + isum[q];
    or_dss(g, 1, u[k].data());
  }
  {
    std::vector<double *> up = ptrs(u);
    constrain2sphere(g, up);
  }
  rhs(u, K[0]);
  for (int k = 0; k < nev; k++) or_dss(g, 1, K[0][k].data());
  for (int k = 0; k < nev; k++) {
    double cm = c[nstage - 1] * h;
    OR_PFOR
    for (size_t q = 0; q < n; q++) isum[q] = K[0][k][q] * cm;
    OR_PFOR
    for (size_t q = 0; q < n; q++) uout[k][q] += isum[q];
    or_dss(g, 1, uout[k]);
  }
  {
    std::vector<double *> up(uout, uout + nev);
    constrain2sphere(g, up);  // step_b end
    constrain2sphere(g, up);  // GBurgers::step_exrk (gburgers.ipp:405)
  }
  (void)t;
}

// ---------------------------------------------------------------------------
// Adjoint of the Burgers RHS and RK step (no reference counterpart; the
// transposes of the operators above, for the FP64 CPU reference gradient).
//   F_c(u) = imass * (-nu L u_c - M * sum_j u_j * d_j u_c)
//   J_F^T w:  z = imass w_c;  lam_c += -nu L z - sum_j d_j^T(u_j M z);
//             lam_j += -(d_j u_c) M z
// with d_j^T v = D_j^T(dXi_j/dX_j v) on regular grids.  Regular grids only.
// ---------------------------------------------------------------------------
void or_burgers_rhs_adjoint(OrGrid *g, double nu, int nstate, const double *const *u,
                            const double *const *w, double *const *lam) {
  size_t n = g->ndof;
  std::vector<double> z(n), Lz(n), mz(n), t(n), v(n), dt(n);
  double one = 1.0;
  for (int c = 0; c < nstate; c++) {
    for (size_t q = 0; q < n; q++) z[q] = g->imass[q] * w[c][q];
    or_helmholtz_apply(g, &one, 1, nullptr, z.data(), Lz.data());
    for (size_t q = 0; q < n; q++) {
      lam[c][q] -= nu * Lz[q];
      mz[q] = g->mass[q] * z[q];
    }
    for (int j = 0; j < nstate; j++) {
      or_deriv(g, u[c], j + 1, t.data());
      for (size_t q = 0; q < n; q++) lam[j][q] -= t[q] * mz[q];
      for (size_t q = 0; q < n; q++) v[q] = g->dxidx[j][0][q] * (u[j][q] * mz[q]);
      or_grefderiv(g, v.data(), j + 1, 1, dt.data());
      for (size_t q = 0; q < n; q++) lam[c][q] -= dt[q];
    }
  }
}

// lam: in dJ/du_{n+1}, out dJ/du_n for one or_burgers_step (mode 0) from un.
void or_burgers_step_adjoint(OrGrid *g, double nu, int order, int nstate, double dt,
                             const double *const *un, double *const *lam) {
  size_t n = g->ndof;
  std::vector<double> alpha, beta, c;
  butcher(order, alpha, beta, c);
  double h = dt;
  int ns = order;
  typedef std::vector<std::vector<double>> Fld;
  // forward: stage inputs us[m] and smoothed stage derivatives K[m]
  std::vector<Fld> us(ns, Fld(nstate, std::vector<double>(n))), K(ns, Fld(nstate, std::vector<double>(n)));
  for (int m = 0; m < ns; m++) {
    for (int k = 0; k < nstate; k++) {
      for (size_t q = 0; q < n; q++) {
        double isum = 0.0;
        for (int j = 0; j < m; j++) isum = 1.0 * isum + beta[(size_t)m * order + j] * h * K[j][k][q];
        us[m][k][q] = un[k][q] + isum;
      }
      if (m == ns - 1) or_dss(g, 1, us[m][k].data());
    }
    std::vector<const double *> up(nstate);
    std::vector<double *> kp(nstate);
    for (int k = 0; k < nstate; k++) {
      up[k] = us[m][k].data();
      kp[k] = K[m][k].data();
    }
    or_burgers_rhs(g, 0, nu, nstate, up.data(), kp.data());
    for (int k = 0; k < nstate; k++) or_dss(g, 1, K[m][k].data());
  }
  // reverse
  Fld lin(nstate, std::vector<double>(n)), ab(nstate, std::vector<double>(n));
  std::vector<Fld> kb(ns, Fld(nstate, std::vector<double>(n, 0.0)));
  for (int k = 0; k < nstate; k++) {
    std::copy(lam[k], lam[k] + n, ab[k].begin());
    or_dss(g, 1, ab[k].data());
    lin[k] = ab[k];  // uout starts as uin
    for (int m = 0; m < ns; m++)
      for (size_t q = 0; q < n; q++) kb[m][k][q] += c[m] * h * ab[k][q];
  }
  for (int m = ns - 1; m >= 0; m--) {
    Fld y(nstate, std::vector<double>(n, 0.0));
    std::vector<const double *> up(nstate), wp(nstate);
    std::vector<double *> yp(nstate);
    for (int k = 0; k < nstate; k++) {
      or_dss(g, 1, kb[m][k].data());
      up[k] = us[m][k].data();
      wp[k] = kb[m][k].data();
      yp[k] = y[k].data();
    }
    or_burgers_rhs_adjoint(g, nu, nstate, up.data(), wp.data(), yp.data());
    for (int k = 0; k < nstate; k++) {
      if (m == ns - 1) or_dss(g, 1, y[k].data());
      for (size_t q = 0; q < n; q++) lin[k][q] += y[k][q];
      for (int j = 0; j < m; j++)
        for (size_t q = 0; q < n; q++) kb[j][k][q] += beta[(size_t)m * order + j] * h * y[k][q];
    }
  }
  for (int k = 0; k < nstate; k++) std::copy(lin[k].begin(), lin[k].end(), lam[k]);
}

// GCG::solve_impl (gcg.ipp:239-345, 361-382) and compute_norm (395-429)
int or_cg_solve(OrGrid *g, const double *p, size_t np, const double *q, const double *b,
                const double *xb, double *x, const OrCGTraits *tr, int *iters, double *resid) {
  size_t n = g->ndof;
  bool bbv = xb != nullptr;
  std::vector<double> qv(n), r(n), w(n), z(n), imult(n), dinv;
  or_imult(g, imult.data());
  if (tr->precond == 1) {
    dinv.resize(n);
    or_helmholtz_diag_inv(g, p, np, q, dinv.data());
  }
  auto gdot = [&](const std::vector<double> &a, const std::vector<double> &bv) {
    double s = 0.0;
    for (size_t j = 0; j < n; j++) s += a[j] * bv[j] * imult[j];
    return s;
  };
  auto norm = [&](const double *v) {
    switch (tr->normtype) {
      case 0: {
        double m = 0.0;
        for (size_t j = 0; j < n; j++) m = std::max(m, std::fabs(v[j]));
        return m;
      }
      case 1: {
        double s = 0.0;
        for (size_t j = 0; j < n; j++) s += v[j] * v[j];
        return std::sqrt(s / (double)n);
      }
      case 2:
      case 3: {
        std::vector<double> t(n);
        for (size_t j = 0; j < n; j++) t[j] = tr->normtype == 2 ? v[j] * v[j] : std::fabs(v[j]);
        double s = or_integrate(g, t.data()) * (1.0 / g->volume);
        return tr->normtype == 2 ? std::sqrt(s) : s;
      }
      default: {
        double s = 0.0;
        for (size_t j = 0; j < n; j++) s += v[j] * v[j] * imult[j];
        return std::sqrt(s);
      }
    }
  };
  auto apply = [&](const double *in, double *out) { or_helmholtz_apply(g, p, np, q, in, out); };
  if (bbv) {
    for (size_t j = 0; j < n; j++) x[j] *= g->mask[j];
    for (size_t j = 0; j < n; j++) x[j] += xb[j];
  }
  int iret = 0;
  std::copy(b, b + n, r.begin());
  apply(x, w.data());
  for (size_t j = 0; j < n; j++) r[j] -= w[j];
  or_dss(g, 0, r.data());
  if (bbv)
    for (size_t j = 0; j < n; j++) r[j] *= g->mask[j];
  if (tr->precond == 1)
    for (size_t j = 0; j < n; j++) z[j] = dinv[j] * r[j];
  else
    z = r;
  w = z;
  double rho = gdot(r, z);
  double rnorm, rtol;
  if (tr->normtype == 4) {
    std::vector<double> ba(b, b + n);
    or_dss(g, 0, ba.data());
    if (bbv)
      for (size_t j = 0; j < n; j++) ba[j] *= g->mask[j];
    rnorm = norm(ba.data());
    rtol = tr->tol * rnorm;
  } else {
    rnorm = norm(b);
    rtol = rnorm < 1.0 ? tr->tol : tr->tol * rnorm;
  }
  int it = 0;
  rnorm = 10.0 * rtol;
  while (iret == 0 && it < tr->maxit && rnorm > rtol) {
    apply(w.data(), qv.data());
    or_dss(g, 0, qv.data());
    if (bbv)
      for (size_t j = 0; j < n; j++) qv[j] *= g->mask[j];
    double alpha = rho / gdot(w, qv);
    for (size_t j = 0; j < n; j++) x[j] = 1.0 * x[j] + alpha * w[j];
    for (size_t j = 0; j < n; j++) r[j] = 1.0 * r[j] + (-alpha) * qv[j];
    if (tr->precond == 1)
      for (size_t j = 0; j < n; j++) z[j] = dinv[j] * r[j];
    else
      z = r;
    double rhom = rho;
    rho = gdot(r, z);
    double beta = rho / rhom;
    for (size_t j = 0; j < n; j++) w[j] = beta * w[j] + 1.0 * z[j];
    rnorm = norm(r.data());
    if (resid) resid[it] = rnorm;
    it++;
  }
  if (bbv)
    for (size_t j = 0; j < n; j++) x[j] *= g->mask[j];
  if (iret == 0 && it >= tr->maxit && rnorm > tr->tol) iret = 1;  // GCGERR_NOCONVERGE
  if (bbv)
    for (size_t j = 0; j < n; j++) x[j] += xb[j];
  if (iters) *iters = it;
  return iret;
}

}  // extern "C"

namespace {
// Common tail of grid construction: GGFX maps, mass (GMass::init3d/2d,
// gmass.ipp:134-212), inverse mass via DSS-Smooth (gmass.ipp:75-82),
// multiplicity and volume (ggrid.ipp:514-521).
void finish_grid(OrGrid &g, double tol_fac) {
  ggfx_match(g, tol_fac * g.minnodedist);
  size_t n = g.ndof;
  g.buf.assign(n, std::vector<size_t>());
  for (size_t lid = 0; lid < n; lid++)
    for (size_t id : g.group[lid]) g.buf[id].push_back(lid);
  g.mass.assign(n, 0.0);
  size_t q = 0;
  for (size_t e = 0; e < g.nelems; e++) {
    if (g.dim == 2) {
      for (int k = 0; k < g.N[1]; k++)
        for (int j = 0; j < g.N[0]; j++, q++) g.mass[q] = g.basis[1].w[k] * g.basis[0].w[j];
    } else {
      for (int l = 0; l < g.N[2]; l++)
        for (int k = 0; k < g.N[1]; k++)
          for (int j = 0; j < g.N[0]; j++, q++)
            g.mass[q] = g.basis[2].w[l] * g.basis[1].w[k] * g.basis[0].w[j];
    }
  }
  for (size_t j = 0; j < n; j++) g.mass[j] *= g.jac[j];
  g.imass = g.mass;
  or_dss(&g, 1, g.imass.data());
  for (size_t j = 0; j < n; j++) g.imass[j] = 1.0 / g.imass[j];
  g.imult.resize(n);
  or_imult(&g, g.imult.data());
  std::vector<double> one(n, 1.0);
  g.volume = or_integrate(&g, one.data());
}
}  // namespace

// Embedded-grid metric G(i,j) = sum_k dXidX(i,k) dXidX(j,k) W0 W1 Jac
// (def_init 2-D branch, ghelmholtz.ipp:476-491), computed on demand.
void or_embed_G(OrGrid *g, int i, int j, std::vector<double> &G) {
  size_t n = g->ndof;
  G.assign(n, 0.0);
  if (g->dim == 3) {
    // def_init 3-D (ghelmholtz.ipp:494-514): passes over all (ii, jj) in
    // (jj outer, ii inner) order, with G(ii,jj) and G(jj,ii) one vector, so
    // the off-diagonal vector of (i, j) takes its terms in the (min, max) pass
    // order of both members.
    auto same = [&](int a, int b) { return (a == i && b == j) || (a == j && b == i); };
    for (int jj = 0; jj < 3; jj++)
      for (int ii = 0; ii < 3; ii++) {
        if (!same(ii, jj)) continue;
        size_t q = 0;
        for (size_t e = 0; e < g->nelems; e++)
          for (int pp = 0; pp < g->N[2]; pp++)
            for (int m = 0; m < g->N[1]; m++)
              for (int l = 0; l < g->N[0]; l++, q++)
                for (int k = 0; k < 3; k++)
                  G[q] += g->dxidx[ii][k][q] * g->dxidx[jj][k][q] * g->basis[0].w[l] *
                          g->basis[1].w[m] * g->basis[2].w[pp] * g->jac[q];
      }
    return;
  }
  int nxy = g->nxy();
  size_t q = 0;
  for (size_t e = 0; e < g->nelems; e++)
    for (int m = 0; m < g->N[1]; m++)
      for (int l = 0; l < g->N[0]; l++, q++)
        for (int k = 0; k < nxy; k++)
          G[q] += g->dxidx[i][k][q] * g->dxidx[j][k][q] * g->basis[0].w[l] * g->basis[1].w[m] * g->jac[q];
}

// ---------------------------------------------------------------------------
// Icosahedral sphere grid: GGridIcos 2-D (src/cdg/grid/ggrid_icos.ipp)
// ---------------------------------------------------------------------------
namespace icos {

const double PI_ = 3.14159265358979323846264338328;  // gtypes.h:92

struct Pt {  // GTPoint<double>, 3 components
  double x1 = 0, x2 = 0, x3 = 0;
  Pt() {}
  Pt(double a, double b, double c) : x1(a), x2(b), x3(c) {}
  Pt operator+(const Pt &o) const { return Pt(x1 + o.x1, x2 + o.x2, x3 + o.x3); }
  Pt operator-(const Pt &o) const { return Pt(x1 - o.x1, x2 - o.x2, x3 - o.x3); }
  Pt operator*(double s) const { return Pt(x1 * s, x2 * s, x3 * s); }
  double norm() const {  // gtpoint.hpp:72-74
    double xn = 0;
    xn += x1 * x1;
    xn += x2 * x2;
    xn += x3 * x3;
    return std::sqrt(xn);
  }
};

// GTVector::sortincreasing (2) via quicksorts2l/partitions2l
// (gtvector.ipp:2965-2999, 3052-3090, 3229-3243): sorts a in place.
void partition_sort(std::vector<double> &a, std::vector<size_t> &is, long start, long end) {
  if (start >= end) return;
  long P = start + (end - start) / 2;
  double pivot = a[P];
  std::swap(a[P], a[start]);
  std::swap(is[P], is[start]);
  long i = start + 1, j = end;
  while (i <= j) {
    while (i <= j && a[i] <= pivot) i++;
    while (i <= j && a[j] > pivot) j--;
    if (i < j) {
      std::swap(a[i], a[j]);
      std::swap(is[i], is[j]);
    }
  }
  std::swap(a[i - 1], a[start]);
  std::swap(is[i - 1], is[start]);
  partition_sort(a, is, start, i - 2);
  partition_sort(a, is, i, end);
}
void sortincreasing(std::vector<double> &a, std::vector<size_t> &is) {
  is.resize(a.size());
  for (size_t j = 0; j < a.size(); j++) is[j] = j;
  partition_sort(a, is, 0, (long)a.size() - 1);
}

// lagvert (ggrid_icos.ipp:2140-2173)
void lagvert(const Pt &a, const Pt &b, const Pt &c, long I, long nrows, std::vector<Pt> &R) {
  R.resize(I + 1);
  double fact = 1.0 / static_cast<double>(nrows + 1);
  double xI = static_cast<double>(I);
  Pt rL = a + ((b - a) * (xI * fact));
  Pt rR = a + ((c - a) * (xI * fact));
  fact = I > 0 ? 1.0 / static_cast<double>(I) : 1.0;
  for (long j = 0; j < I + 1; j++) {
    double xJ = static_cast<double>(j);
    R[j] = rL + (rR - rL) * (xJ * fact);
  }
}

// project2sphere (1)/(2) (ggrid_icos.ipp:1459-1515)
Pt proj(const Pt &v, double rad) {
  double r = v.norm();
  double xlat = std::asin(v.x3 / r);
  double xlong = std::atan2(v.x2, v.x1);
  xlong = xlong < 0.0 ? 2.0 * PI_ + xlong : xlong;
  return Pt(rad * std::cos(xlat) * std::cos(xlong), rad * std::cos(xlat) * std::sin(xlong),
            rad * std::sin(xlat));
}

// xyz2spherical (3) (ggrid_icos.ipp:1700-1720): (r, lat, long)
Pt sph(const Pt &p) {
  double x = p.x1, y = p.x2, z = p.x3;
  double r = std::sqrt(x * x + y * y + z * z);
  double lon = std::atan2(y, x);
  lon = lon < 0.0 ? 2.0 * PI_ + lon : lon;
  return Pt(r, std::asin(z / r), lon);
}

// GLLBasis::evalBasis (2), evalDBasis (4) (gllbasis.ipp:1300-1327, 1628-1665)
double evalBasis(const Basis &b, int i, double eta) {
  LD ttiny = 100.0L * std::numeric_limits<LD>::epsilon();
  LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  int Np = b.p;
  LD fact = -1.0 / (Np * (Np + 1.0)), xi = static_cast<LD>(eta);
  double fRet = 1.0;
  if (xi < -1.0L || xi > 1.0L) fRet = 0.0;
  else if (std::fabs(xi - b.xl[i]) > tetiny) {
    LD ppn_i = b.Pnl[i], ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2;
    jacobi(Np, 0.0L, 0.0L, ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2, xi);
    LD gfact = (1.0L - xi * xi) / (xi - b.xl[i] + ttiny);
    fRet = static_cast<double>(fact * gfact * pder_xi / (ppn_i + ttiny));
  }
  return fRet;
}
double evalDBasis(const Basis &b, int j, double eta) {
  LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  int Np = b.p;
  LD fact = -1.0 / (Np * (Np + 1.0)), xi = static_cast<LD>(eta);
  LD g1 = xi - b.xl[j];
  double fRet = 0.0;
  if (xi == -1.0L && j == 0) fRet = -0.25 * Np * (Np + 1.0);
  else if (xi == 1.0L && j == Np) fRet = 0.25 * Np * (Np + 1.0);
  else if (std::fabs(g1) > tetiny) {
    LD ppn_j = b.Pnl[j], ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2;
    jacobi(Np, 0.0L, 0.0L, ppn_xi, pder_xi, pm1, pdm1, pm2, pdm2, xi);
    LD gfact = fact / ppn_j, g1i = 1.0L / g1, g2 = (1.0L - xi * xi) * g1i;
    LD pdd = 2.0L * xi * pder_xi - Np * (Np + 1.) * ppn_xi;
    fRet = static_cast<double>(gfact * g1i * (pdd - (2.0L * xi + g2) * pder_xi));
  }
  return fRet;
}

}  // namespace icos

// GGridIcos::init2d/lagrefine/order_triangles/do_elems2d on one rank, with
// GElem_base::dogeom2d for GE_2DEMBEDDED (gelem_base.cpp:877-955) using the
// embedded shape functions (gshapefcn_embed.ipp:185-205, 356-390), Jac
// (gelem_base.cpp:1066-1170) and inv (:1249-1300); GDD_base::doDD on one
// rank keeps all triangles in order (gdd_base.ipp:73-102).  refine: 0 for
// GICOS_BISECTION (nrows = 2^ilevel - 1), 1 for GICOS_LAGRANGIAN.
extern "C" OrGrid *or_grid_icos_create2(int p, int ilevel, int refine, double radius,
                                        double tol_fac) {
  using namespace icos;
  if (p < 1 || ilevel < 0 || (refine != 0 && refine != 1)) return nullptr;
  OrGrid *gp = new OrGrid();
  OrGrid &g = *gp;
  g.dim = 2;
  g.gtype = 2;
  if (!gll(p, g.basis[0]) || !gll(p, g.basis[1])) {
    delete gp;
    return nullptr;
  }
  g.N[0] = g.N[1] = p + 1;
  g.N[2] = 1;
  g.nnodes = (size_t)g.N[0] * g.N[1];

  // init2d (ggrid_icos.ipp:139-230)
  double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  double fv0[12][3] = {{0.000000000000000, 0.000000000000000, 1.000000000000000},
                       {0.894427190999916, 0.000000000000000, 0.447213595499958},
                       {-0.894427190999916, 0.000000000000000, -0.447213595499958},
                       {0.000000000000000, 0.000000000000000, -1.000000000000000},
                       {-0.723606797749979, 0.525731112119134, 0.447213595499958},
                       {-0.723606797749979, -0.525731112119134, 0.447213595499958},
                       {0.723606797749979, 0.525731112119134, -0.447213595499958},
                       {0.723606797749979, -0.525731112119134, -0.447213595499958},
                       {0.276393202250021, 0.850650808352040, 0.447213595499958},
                       {0.276393202250021, -0.850650808352040, 0.447213595499958},
                       {-0.276393202250021, 0.850650808352040, -0.447213595499958},
                       {-0.276393202250021, -0.850650808352040, -0.447213595499958}};
  double fact = 1.0 / std::sqrt(phi * phi + 1.0);
  for (auto &r : fv0)
    for (double &v : r) v *= fact;
  int ifv0[20][3] = {{0, 1, 8},  {0, 8, 4},   {0, 4, 5},  {0, 5, 9},  {0, 9, 1},
                     {1, 6, 8},  {8, 6, 10},  {8, 10, 4}, {4, 10, 2}, {4, 2, 5},
                     {5, 2, 11}, {5, 11, 9},  {9, 11, 7}, {9, 7, 1},  {1, 7, 6},
                     {3, 6, 7},  {3, 7, 11},  {3, 11, 2}, {3, 2, 10}, {3, 10, 6}};

  // lagrefine (ggrid_icos.ipp:254-328)
  long nrows = refine == 0 ? (long)std::pow(2, ilevel) - 1 : ilevel;
  std::vector<std::array<Pt, 3>> tmesh(20 * (nrows * (nrows + 2) + 1));
  size_t n = 0;
  std::vector<Pt> R0, R1, Rz(2 * (nrows + 1) + 1);
  for (int t = 0; t < 20; t++) {
    Pt a(fv0[ifv0[t][0]][0], fv0[ifv0[t][0]][1], fv0[ifv0[t][0]][2]);
    Pt b(fv0[ifv0[t][1]][0], fv0[ifv0[t][1]][1], fv0[ifv0[t][1]][2]);
    Pt c(fv0[ifv0[t][2]][0], fv0[ifv0[t][2]][1], fv0[ifv0[t][2]][2]);
    for (long l = 0; l < nrows + 1; l++) {
      lagvert(a, b, c, l, nrows, R0);
      lagvert(a, b, c, l + 1, nrows, R1);
      for (long j = 0; j < 2 * l + 3; j++) Rz[j] = (j % 2 == 0) ? R1[j / 2] : R0[(j - 1) / 2];
      for (long m = 0; m < 2 * l + 1; m++) tmesh[n++] = {Rz[m], Rz[m + 1], Rz[m + 2]};
    }
  }
  for (auto &tr : tmesh)
    for (auto &v : tr) v = proj(v, 1.0);
  // order_triangles (ggrid_icos.ipp:2202-2258)
  std::vector<int> iup(tmesh.size(), 0);
  for (size_t i = 0; i < tmesh.size(); i++) {
    std::array<Pt, 3> cverts = tmesh[i];
    std::vector<double> lat(3), lon(3);
    for (int j = 0; j < 3; j++) {
      Pt s = sph(cverts[j]);
      lat[j] = s.x2;
      lon[j] = s.x3;
    }
    std::vector<size_t> isl;
    sortincreasing(lon, isl);
    if (std::fabs(lon[isl[0]] - lon[isl[2]]) < PI_)
      for (int j = 0; j < 3; j++)
        if (lon[j] > 1.5 * PI_ && lon[j] <= 2.0 * PI_) lon[j] -= 2.0 * PI_;
    sortincreasing(lon, isl);
    if (lat[isl[1]] > lat[isl[0]] && lat[isl[1]] > lat[isl[2]]) {
      tmesh[i] = {cverts[isl[0]], cverts[isl[2]], cverts[isl[1]]};
      iup[i] = 1;
    }
    if (lat[isl[1]] < lat[isl[0]] && lat[isl[1]] < lat[isl[2]])
      tmesh[i] = {cverts[isl[1]], cverts[isl[2]], cverts[isl[0]]};
  }

  // do_elems2d (ggrid_icos.ipp:379-513)
  const Basis &B0 = g.basis[0], &B1 = g.basis[1];
  int N0 = g.N[0], N1 = g.N[1];
  size_t nn = g.nnodes;
  g.nelems = 3 * tmesh.size();
  g.ndof = g.nelems * nn;
  for (int d = 0; d < 3; d++) {
    g.xn[d].assign(g.ndof, 0.0);
    for (int j = 0; j < 3; j++) g.dxidx[d][j].assign(g.ndof, 0.0);
  }
  g.jac.assign(g.ndof, 0.0);
  std::vector<double> Ni(nn), xg[3], dNi(nn), tmp(nn);
  std::vector<double> rij[3][3];
  // h_b(xi_q) and h_b'(xi_q) tables, [b*N + q]
  std::vector<double> h0((size_t)N0 * N0), dh0((size_t)N0 * N0), h1((size_t)N1 * N1),
      dh1((size_t)N1 * N1);
  for (int b = 0; b < N0; b++)
    for (int q = 0; q < N0; q++) {
      h0[(size_t)b * N0 + q] = evalBasis(B0, b, B0.xi[q]);
      dh0[(size_t)b * N0 + q] = evalDBasis(B0, b, B0.xi[q]);
    }
  for (int b = 0; b < N1; b++)
    for (int q = 0; q < N1; q++) {
      h1[(size_t)b * N1 + q] = evalBasis(B1, b, B1.xi[q]);
      dh1[(size_t)b * N1 + q] = evalDBasis(B1, b, B1.xi[q]);
    }
  for (size_t i = 0, icurr = 0; i < tmesh.size(); i++) {
    Pt v1 = tmesh[i][0], v2 = tmesh[i][1], v3 = tmesh[i][2];
    double third = 1.0 / 3.0;
    Pt ct = (v1 + (v2 + v3)) * third;
    for (int j = 0; j < 3; j++, icurr += nn) {
      Pt cv[4];
      if (iup[i] == 1) {
        if (j == 0) { cv[0] = v1; cv[1] = (v1 + v2) * 0.5; cv[2] = ct; cv[3] = (v1 + v3) * 0.5; }
        if (j == 1) { cv[0] = (v1 + v2) * 0.5; cv[1] = v2; cv[2] = (v2 + v3) * 0.5; cv[3] = ct; }
        if (j == 2) { cv[0] = ct; cv[1] = (v2 + v3) * 0.5; cv[2] = v3; cv[3] = (v1 + v3) * 0.5; }
      } else {
        if (j == 0) { cv[0] = v1; cv[1] = (v1 + v2) * 0.5; cv[2] = ct; cv[3] = (v1 + v3) * 0.5; }
        if (j == 1) { cv[0] = ct; cv[1] = (v1 + v2) * 0.5; cv[2] = v2; cv[3] = (v2 + v3) * 0.5; }
        if (j == 2) { cv[0] = (v1 + v3) * 0.5; cv[1] = ct; cv[2] = (v2 + v3) * 0.5; cv[3] = v3; }
      }
      for (auto &v : cv) v = proj(v, 1.0);
      // reorderverts2d with cart2gnomonic (ggrid_icos.ipp:1733-1775, 1923-1988)
      Pt c = (cv[0] + cv[1] + cv[2] + cv[3]) * 0.25;
      double xlatc = std::asin(c.x3), xlongc = std::atan2(c.x2, c.x1);
      xlongc = xlongc < 0.0 ? 2 * PI_ + xlongc : xlongc;
      double eps = 10.0 * std::numeric_limits<double>::epsilon();
      Pt tv[4];
      for (int m = 0; m < 4; m++) {
        double r = cv[m].norm();
        double xlat = std::asin(cv[m].x3 / r), xlong = std::atan2(cv[m].x2, cv[m].x1);
        xlong = xlong < 0.0 ? 2.0 * PI_ + xlong : xlong;
        double den = std::sin(xlatc) * std::sin(xlat) +
                     std::cos(xlatc) * std::cos(xlat) * std::cos(xlong - xlongc);
        den = std::fabs(den) < eps ? 0.0 : 1.0 / den;
        tv[m].x1 = 1.0 * std::cos(xlat) * std::sin(xlong - xlongc) * den;
        tv[m].x2 = 1.0 * (std::cos(xlatc) * std::sin(xlat) -
                          std::sin(xlatc) * std::cos(xlat) * std::cos(xlong - xlongc)) * den;
      }
      std::vector<double> xs = {tv[0].x1, tv[1].x1, tv[2].x1, tv[3].x1};
      std::vector<size_t> Ixy;
      sortincreasing(xs, Ixy);
      size_t is[4];
      if (tv[Ixy[0]].x2 < tv[Ixy[1]].x2) { is[0] = Ixy[0]; is[3] = Ixy[1]; }
      else { is[0] = Ixy[1]; is[3] = Ixy[0]; }
      if (tv[Ixy[2]].x2 < tv[Ixy[3]].x2) { is[1] = Ixy[2]; is[2] = Ixy[3]; }
      else { is[1] = Ixy[3]; is[2] = Ixy[2]; }
      Pt gv[4];
      for (int m = 0; m < 4; m++) gv[m] = cv[is[m]];
      // node coordinates: sum_m Ni_m * (v_m * 0.25) (GShapeFcn_linear::Ni_2d,
      // gshapefcn_linear.ipp:171-216), then project2sphere (3) to radius
      for (int l = 0; l < 3; l++) {
        xg[l].assign(nn, 0.0);
        for (int m = 0; m < 4; m++) {
          size_t q = 0;
          for (int jj = 0; jj < N1; jj++) {
            double sm = 1.0 - B1.xi[jj], sp = 1.0 + B1.xi[jj];
            for (int ii = 0; ii < N0; ii++, q++) {
              double rm = 1.0 - B0.xi[ii], rp = 1.0 + B0.xi[ii];
              Ni[q] = m == 0 ? rm * sm : m == 1 ? rp * sm : m == 2 ? rp * sp : rm * sp;
            }
          }
          double vml = l == 0 ? gv[m].x1 : l == 1 ? gv[m].x2 : gv[m].x3;
          for (size_t q2 = 0; q2 < nn; q2++) xg[l][q2] += Ni[q2] * (vml * 0.25);
        }
      }
      for (size_t q = 0; q < nn; q++) {
        double x = xg[0][q], y = xg[1][q], z = xg[2][q];
        double r = std::sqrt(x * x + y * y + z * z);
        double xlat = std::asin(z / r), xlong = std::atan2(y, x);
        xlong = xlong < 0.0 ? 2.0 * PI_ + xlong : xlong;
        g.xn[0][icurr + q] = radius * std::cos(xlat) * std::cos(xlong);
        g.xn[1][icurr + q] = radius * std::cos(xlat) * std::sin(xlong);
        g.xn[2][icurr + q] = radius * std::sin(xlat);
      }
      // dogeom2d, embedded: rij(k,m) = sum_I dN_I/dxi^m x^k_I
      for (int m = 0; m < 3; m++)
        for (int k = 0; k < 3; k++) {
          tmp.assign(nn, 0.0);
          for (int jb = 0, nI = 0; jb < N1; jb++)
            for (int ib = 0; ib < N0; ib++, nI++) {
              // GShapeFcn_embed::dNdXi_2d (gshapefcn_embed.ipp:356-390)
              const double *d0 = (m == 0 ? dh0 : h0).data() + (size_t)ib * N0;
              const double *d1 = (m == 1 ? dh1 : h1).data() + (size_t)jb * N1;
              for (int qj = 0, q = 0; qj < N1; qj++)
                for (int qi = 0; qi < N0; qi++, q++) dNi[q] = d0[qi] * d1[qj];
              double xk = g.xn[k][icurr + nI];
              for (size_t q = 0; q < nn; q++) {
                dNi[q] *= xk;
                tmp[q] += dNi[q];
              }
            }
          rij[k][m] = tmp;
        }
      for (size_t q = 0; q < nn; q++) {
        auto G = [&](int a, int b) { return rij[a][b][q]; };
        double det = G(0, 0) * (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) -
                     G(0, 1) * (G(1, 0) * G(2, 2) - G(1, 2) * G(2, 0)) +
                     G(0, 2) * (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0));
        double jv = std::fabs(det);
        double x = G(0, 2), y = G(1, 2), z = G(2, 2);
        jv /= std::sqrt(x * x + y * y + z * z);
        g.jac[icurr + q] = jv;
        double ijac = 1.0 / det;
        size_t id = icurr + q;
        g.dxidx[0][0][id] = (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) * ijac;
        g.dxidx[0][1][id] = -(G(0, 1) * G(2, 2) - G(2, 1) * G(0, 2)) * ijac;
        g.dxidx[0][2][id] = (G(0, 1) * G(1, 2) - G(0, 2) * G(1, 1)) * ijac;
        g.dxidx[1][0][id] = -(G(1, 0) * G(2, 2) - G(2, 0) * G(1, 2)) * ijac;
        g.dxidx[1][1][id] = (G(0, 0) * G(2, 2) - G(2, 0) * G(0, 2)) * ijac;
        g.dxidx[1][2][id] = -(G(0, 0) * G(1, 2) - G(0, 2) * G(1, 0)) * ijac;
        g.dxidx[2][0][id] = (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0)) * ijac;
        g.dxidx[2][1][id] = -(G(0, 0) * G(2, 1) - G(0, 1) * G(2, 0)) * ijac;
        g.dxidx[2][2][id] = (G(0, 0) * G(1, 1) - G(1, 0) * G(0, 1)) * ijac;
      }
    }
  }
  // find_min_dist (ggrid.ipp:977-1013), no periodicity, no boundaries
  double tiny = 1000.0 * std::numeric_limits<double>::epsilon(), dmin = std::numeric_limits<double>::max();
  for (size_t q = 0; q + 1 < g.ndof; q++) {
    double dx = g.xn[0][q + 1] - g.xn[0][q], dy = g.xn[1][q + 1] - g.xn[1][q],
           dz = g.xn[2][q + 1] - g.xn[2][q];
    double r = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (r > tiny) dmin = std::min(dmin, r);
  }
  g.minnodedist = dmin;
  for (int d = 0; d < 3; d++) g.xp[d] = g.xn[d];
  g.mask.assign(g.ndof, 1.0);
  finish_grid(g, tol_fac);
  return gp;
}

extern "C" OrGrid *or_grid_icos_create(int p, int ilevel, double radius, double tol_fac) {
  return or_grid_icos_create2(p, ilevel, 0, radius, tol_fac);
}

namespace {
// GElem_base::dogeom2d for GE_DEFORMED (gelem_base.cpp:877-955) from g.xn:
// rij(k,m) = sum_I dN_I/dxi^m x^k_I with the tensor-product GLL shape
// functions, Jac = rij00 rij11 - rij01 rij10 (GElem_base::Jac, 2x2,
// gelem_base.cpp:1121-1136) and the 2x2 branch of GElem_base::inv
// (gelem_base.cpp:1255-1266), whose off-diagonal entries are -rij10/jac at
// (0,1) and -rij01/jac at (1,0); ref_inverse = 0 swaps those two (the true
// inverse).
void deformed_geom2d(OrGrid &g, int ref_inverse) {
  using namespace icos;
  const int N = g.N[0];
  const size_t nn = (size_t)N * N;
  const Basis &B = g.basis[0];
  std::vector<double> h0((size_t)N * N), dh0((size_t)N * N);  // [basis][node]
  for (int i = 0; i < N; i++)
    for (int q = 0; q < N; q++) {
      h0[(size_t)i * N + q] = evalBasis(B, i, B.xi[q]);
      dh0[(size_t)i * N + q] = evalDBasis(B, i, B.xi[q]);
    }
  std::vector<double> dNi(nn), tmp(nn), rij[2][2];
  for (size_t e = 0; e < g.nelems; e++) {
    const size_t b0 = e * nn;
    for (int m = 0; m < 2; m++)
      for (int k = 0; k < 2; k++) {
        tmp.assign(nn, 0.0);
        for (int jb = 0, nI = 0; jb < N; jb++)
          for (int ib = 0; ib < N; ib++, nI++) {
            // dN_I/dxi^m at every node: h'_ib(xi) h_jb(eta) or h_ib(xi) h'_jb(eta)
            const double *d0 = (m == 0 ? dh0 : h0).data() + (size_t)ib * N;
            const double *d1 = (m == 1 ? dh0 : h0).data() + (size_t)jb * N;
            for (int qj = 0, q = 0; qj < N; qj++)
              for (int qi = 0; qi < N; qi++, q++) dNi[q] = d0[qi] * d1[qj];
            double xk = g.xn[k][b0 + nI];
            for (size_t q = 0; q < nn; q++) {
              dNi[q] *= xk;
              tmp[q] += dNi[q];
            }
          }
        rij[k][m] = tmp;
      }
    for (size_t q = 0; q < nn; q++) {
      auto G = [&](int a, int b) { return rij[a][b][q]; };
      double jv = G(0, 0) * G(1, 1) - G(0, 1) * G(1, 0);
      double ijac = 1.0 / (G(0, 0) * G(1, 1) - G(1, 0) * G(0, 1));
      size_t id = b0 + q;
      g.jac[id] = jv;
      g.dxidx[0][0][id] = G(1, 1) * ijac;
      g.dxidx[0][1][id] = -(ref_inverse ? G(1, 0) : G(0, 1)) * ijac;
      g.dxidx[1][0][id] = -(ref_inverse ? G(0, 1) : G(1, 0)) * ijac;
      g.dxidx[1][1][id] = G(0, 0) * ijac;
    }
  }
}
}  // namespace

// 2-D box of GE_DEFORMED quads: the regular box (or_grid_box_create) with
// every node moved by y' = y + h(x) (P1y - y)/H, h(x) = amp (1 - cos(2 pi
// (x - P0x)/Lx))/2, and the deformed geometry of those nodes.  Mask and
// boundary classes stay those of the base box (add_terrain recomputes only
// the geometry, ggrid.ipp:1522-1540).
extern "C" OrGrid *or_grid_terrain_create(int p, const int *ne, const double *P0, const double *P1,
                                          const int *bdy, double amp, int ref_inverse,
                                          double tol_fac) {
  using namespace icos;
  const int pp[2] = {p, p};
  OrGrid *gp = or_grid_box_create(2, pp, ne, P0, P1, bdy, tol_fac);
  if (!gp) return nullptr;
  OrGrid &g = *gp;
  g.gtype = 1;
  const double Lx = P1[0] - P0[0], H = P1[1] - P0[1];
  for (std::vector<double> *c : {g.xn, g.xp})
    for (size_t q = 0; q < g.ndof; q++) {
      double h = amp * 0.5 * (1.0 - std::cos(2.0 * PI_ * (c[0][q] - P0[0]) / Lx));
      c[1][q] = c[1][q] + h * (P1[1] - c[1][q]) / H;
    }
  deformed_geom2d(g, ref_inverse);
  g.minnodedist = find_min_dist(g);
  finish_grid(g, tol_fac);
  return gp;
}

namespace {
// GElem_base::dogeom3d for GE_DEFORMED (gelem_base.cpp:972-1048) from g.xn:
// rij(l,m) = sum_I dN_I/dxi^m x^l_I with the tensor-product shape functions
// (factors multiplied x, y, z), Jac = |det rij| (gelem_base.cpp:1148-1168) and
// the 3x3 branch of GElem_base::inv (cofactors over det, gelem_base.cpp:1268-1283).
void deformed_geom3d(OrGrid &g) {
  using namespace icos;
  const int N = g.N[0];
  const size_t nn = (size_t)N * N * N;
  const Basis &B = g.basis[0];
  std::vector<double> h0((size_t)N * N), dh0((size_t)N * N);
  for (int i = 0; i < N; i++)
    for (int q = 0; q < N; q++) {
      h0[(size_t)i * N + q] = evalBasis(B, i, B.xi[q]);
      dh0[(size_t)i * N + q] = evalDBasis(B, i, B.xi[q]);
    }
  std::vector<double> dNi(nn), tmp(nn), rij[3][3];
  for (size_t e = 0; e < g.nelems; e++) {
    const size_t b0 = e * nn;
    for (int m = 0; m < 3; m++)
      for (int l = 0; l < 3; l++) {
        tmp.assign(nn, 0.0);
        for (int kb = 0, nI = 0; kb < N; kb++)
          for (int jb = 0; jb < N; jb++)
            for (int ib = 0; ib < N; ib++, nI++) {
              const double *d0 = (m == 0 ? dh0 : h0).data() + (size_t)ib * N;
              const double *d1 = (m == 1 ? dh0 : h0).data() + (size_t)jb * N;
              const double *d2 = (m == 2 ? dh0 : h0).data() + (size_t)kb * N;
              for (int qk = 0, q = 0; qk < N; qk++)
                for (int qj = 0; qj < N; qj++)
                  for (int qi = 0; qi < N; qi++, q++) dNi[q] = d0[qi] * d1[qj] * d2[qk];
              double xl = g.xn[l][b0 + nI];
              for (size_t q = 0; q < nn; q++) {
                dNi[q] *= xl;
                tmp[q] += dNi[q];
              }
            }
        rij[l][m] = tmp;
      }
    for (size_t q = 0; q < nn; q++) {
      auto G = [&](int a, int b) { return rij[a][b][q]; };
      double det = G(0, 0) * (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) -
                   G(0, 1) * (G(1, 0) * G(2, 2) - G(1, 2) * G(2, 0)) +
                   G(0, 2) * (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0));
      size_t id = b0 + q;
      g.jac[id] = std::fabs(det);
      double ijac = 1.0 / det;
      g.dxidx[0][0][id] = (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) * ijac;
      g.dxidx[0][1][id] = -(G(0, 1) * G(2, 2) - G(2, 1) * G(0, 2)) * ijac;
      g.dxidx[0][2][id] = (G(0, 1) * G(1, 2) - G(0, 2) * G(1, 1)) * ijac;
      g.dxidx[1][0][id] = -(G(1, 0) * G(2, 2) - G(2, 0) * G(1, 2)) * ijac;
      g.dxidx[1][1][id] = (G(0, 0) * G(2, 2) - G(2, 0) * G(0, 2)) * ijac;
      g.dxidx[1][2][id] = -(G(0, 0) * G(1, 2) - G(0, 2) * G(1, 0)) * ijac;
      g.dxidx[2][0][id] = (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0)) * ijac;
      g.dxidx[2][1][id] = -(G(0, 0) * G(2, 1) - G(0, 1) * G(2, 0)) * ijac;
      g.dxidx[2][2][id] = (G(0, 0) * G(1, 1) - G(1, 0) * G(0, 1)) * ijac;
    }
  }
}
}  // namespace

// 3-D box of GE_DEFORMED hexahedra: the regular box with every node moved by
// z' = z + h(x,y)(P1z - z)/H, h = amp (1 - cos(2 pi (x-P0x)/Lx))(1 - cos(2 pi
// (y-P0y)/Ly))/4, and the deformed geometry of those nodes.
extern "C" OrGrid *or_grid_terrain3_create(int p, const int *ne, const double *P0,
                                           const double *P1, const int *bdy, double amp,
                                           double tol_fac) {
  using namespace icos;
  const int pp[3] = {p, p, p};
  OrGrid *gp = or_grid_box_create(3, pp, ne, P0, P1, bdy, tol_fac);
  if (!gp) return nullptr;
  OrGrid &g = *gp;
  g.gtype = 1;
  const double Lx = P1[0] - P0[0], Ly = P1[1] - P0[1], H = P1[2] - P0[2];
  for (std::vector<double> *c : {g.xn, g.xp})
    for (size_t q = 0; q < g.ndof; q++) {
      double h = amp * 0.25 * (1.0 - std::cos(2.0 * PI_ * (c[0][q] - P0[0]) / Lx)) *
                 (1.0 - std::cos(2.0 * PI_ * (c[1][q] - P0[1]) / Ly));
      c[2][q] = c[2][q] + h * (P1[2] - c[2][q]) / H;
    }
  deformed_geom3d(g);
  g.minnodedist = find_min_dist(g);
  finish_grid(g, tol_fac);
  return gp;
}

// The second half of GGrid::add_terrain (ggrid.ipp:1522-1547) on a 2-D
// deformed grid without periodic directions: install new node coordinates,
// recompute the deformed geometry, the minimum node distance and (here) the
// GGFX maps, mass and volume from them.
extern "C" int or_grid_set_coords2d(OrGrid *g, const double *x, const double *y, int ref_inverse,
                                    double tol_fac) {
  if (!g || g->dim != 2 || g->gtype != 1 || !x || !y) return 1;
  std::copy(x, x + g->ndof, g->xn[0].begin());
  std::copy(y, y + g->ndof, g->xn[1].begin());
  g->xp[0] = g->xn[0];
  g->xp[1] = g->xn[1];
  deformed_geom2d(*g, ref_inverse);
  g->minnodedist = find_min_dist(*g);
  finish_grid(*g, tol_fac);
  return 0;
}
