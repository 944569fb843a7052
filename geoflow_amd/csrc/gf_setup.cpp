// gf_setup.cpp -- host-side setup for the MI355X SEM path: GLL basis, box
// grid generation and GGFX coincidence matching.  No device work here.
//
// Built with -ffp-contract=off so that node coordinates and element metrics
// round exactly as the reference's expressions do.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "geoflow_hip.h"

namespace {

typedef long double LD;

// Legendre P_n and P_n' at x by the three-term recurrence.
void legendre(int n, LD x, LD &P, LD &dP) {
  LD p0 = 1.0L, p1 = x, d0 = 0.0L, d1 = 1.0L;
  if (n == 0) {
    P = 1.0L;
    dP = 0.0L;
    return;
  }
  for (int k = 2; k <= n; k++) {
    LD p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (LD)k;
    LD d2 = d0 + (2.0L * k - 1.0L) * p1;
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d2;
  }
  P = p1;
  dP = d1;
}

}  // namespace

extern "C" {

// GLL nodes are the roots of (1-x^2) P_p'(x).  Interior roots are found by
// Newton on q(x) = P_{p+1}(x) - P_{p-1}(x) (proportional to (1-x^2)P_p'),
// seeded with Chebyshev-Gauss-Lobatto points and iterated to long-double
// convergence, then symmetrised.  Weights 2/(p(p+1) P_p(x)^2); derivative
// matrix D(l,j) = P_p(x_l)/(P_p(x_j)(x_l-x_j) + 100 eps_ld) off the diagonal
// and -/+ p(p+1)/4 at the two corners, zero elsewhere on the diagonal -- the
// formulas of GLLBasis::computeWeights/computeDerivMatrix
// (src/cdg/sem/gllbasis.ipp:514-531, 579-607), including the reference's
// 100*eps(long double) shift in the denominator (gllbasis.ipp:592,596).
int gf_gll_basis(int p, double *xi, double *w, double *D) {
  if (p < 1 || p > 64 || !xi || !w || !D) return GF_ERR_ARG;
  int N = p + 1;
  std::vector<LD> x(N), P(N);
  x[0] = -1.0L;
  x[p] = 1.0L;
  const LD pi = 3.14159265358979323846264338327950288L;
  for (int j = 1; j < p; j++) {
    LD xx = -std::cos(pi * (LD)j / (LD)p);
    for (int it = 0; it < 100; it++) {
      LD Pa, dPa, Pb, dPb;
      legendre(p + 1, xx, Pa, dPa);
      legendre(p - 1, xx, Pb, dPb);
      LD f = Pa - Pb, df = dPa - dPb;
      LD dx = f / df;
      xx -= dx;
      if (std::fabs(dx) <= 4.0L * std::numeric_limits<LD>::epsilon() * (1.0L + std::fabs(xx))) break;
    }
    x[j] = xx;
  }
  for (int j = 0; j < N / 2; j++) {  // exact symmetry about 0
    LD s = 0.5L * (x[N - 1 - j] - x[j]);
    x[j] = -s;
    x[N - 1 - j] = s;
  }
  if (N % 2 == 1) x[N / 2] = 0.0L;
  const LD fact = 2.0L / ((LD)p * ((LD)p + 1.0L));
  for (int i = 0; i < N; i++) {
    LD dP;
    legendre(p, x[i], P[i], dP);
    xi[i] = (double)x[i];
    w[i] = (double)(fact / (P[i] * P[i]));
  }
  const LD ttiny = 100.0L * std::numeric_limits<LD>::epsilon();
  for (int j = 0; j < N; j++)
    for (int l = 0; l < N; l++) {
      LD v = 0.0L;
      if (l != j) v = P[l] / (P[j] * (x[l] - x[j]) + ttiny);
      D[l + (size_t)j * N] = (double)v;
    }
  D[0] = (double)(-0.25L * p * (p + 1.0L));
  D[(size_t)p + (size_t)p * N] = (double)(0.25L * p * (p + 1.0L));
  return GF_OK;
}

// Box grid: elements in linear order x fastest (find_rank_subdomain on one
// rank, src/cdg/grid/ggrid_box.ipp:862-978), nodes from the trilinear shape
// function x = sum_m N_m(xi) (v_m * 2^-dim) (do_elems2d/3d,
// ggrid_box.ipp:254-463; gshapefcn_linear.ipp:171-348), GE_REGULAR metrics
// 2/L and prod(L)/2^dim with L taken from vertex differences
// (gelem_base.cpp:877-1048).  Boundary nodes are those whose coordinate is
// within eps = minnodedist/8 of a domain face (FUZZYEQ, gtypes.h:221).
int gf_box_grid(int dim, const int *p, const int *ne, const double *P0, const double *P1,
                const int *bdy, double *const *xyz, double *const *xyzp, double *mask,
                double *elem_dxidx, double *elem_jac, double *minnodedist) {
  if ((dim != 2 && dim != 3) || !p || !ne || !P0 || !P1 || !bdy) return GF_ERR_ARG;
  int N[3] = {1, 1, 1};
  std::vector<double> xi[3], wq[3];
  for (int d = 0; d < dim; d++) {
    if (p[d] < 1 || ne[d] < 1) return GF_ERR_ARG;
    N[d] = p[d] + 1;
    xi[d].resize(N[d]);
    wq[d].resize(N[d]);
    std::vector<double> Dm((size_t)N[d] * N[d]);
    int rc = gf_gll_basis(p[d], xi[d].data(), wq[d].data(), Dm.data());
    if (rc) return rc;
  }
  size_t nn = (size_t)N[0] * N[1] * N[2];
  size_t nel = 1;
  for (int d = 0; d < dim; d++) nel *= (size_t)ne[d];
  size_t ndof = nel * nn;
  double dx[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) dx[d] = (P1[d] - P0[d]) / (double)ne[d];
  std::vector<double> X[3];
  for (int d = 0; d < dim; d++) X[d].resize(ndof);
  // Shape-function values per vertex and node, built once.
  int nv = dim == 3 ? 8 : 4;
  std::vector<double> Nf((size_t)nv * nn);
  for (int m = 0; m < nv; m++) {
    int bx = (m == 1 || m == 2 || m == 5 || m == 6) ? 1 : 0;
    int by = (m == 2 || m == 3 || m == 6 || m == 7) ? 1 : 0;
    int bz = m >= 4 ? 1 : 0;
    size_t q = 0;
    for (int k = 0; k < N[2]; k++)
      for (int j = 0; j < N[1]; j++)
        for (int i = 0; i < N[0]; i++, q++) {
          double r = bx ? 1.0 + xi[0][i] : 1.0 - xi[0][i];
          double s = by ? 1.0 + xi[1][j] : 1.0 - xi[1][j];
          double v = r * s;
          if (dim == 3) {
            double t = bz ? 1.0 + xi[2][k] : 1.0 - xi[2][k];
            v = v * t;
          }
          Nf[(size_t)m * nn + q] = v;
        }
  }
  const double vfac = dim == 3 ? 0.125 : 0.25;
  size_t e = 0;
  int nk = dim == 3 ? ne[2] : 1;
  for (long k = 0; k < nk; k++)
    for (long j = 0; j < ne[1]; j++)
      for (long i = 0; i < ne[0]; i++, e++) {
        double v0[3] = {P0[0] + i * dx[0], P0[1] + j * dx[1], dim == 3 ? P0[2] + k * dx[2] : 0.0};
        double v1[3];  // opposite corner
        for (int d = 0; d < dim; d++) v1[d] = v0[d] + dx[d];
        for (int d = 0; d < dim; d++) {
          double *out = X[d].data() + e * nn;
          for (size_t q = 0; q < nn; q++) {
            double acc = 0.0;
            for (int m = 0; m < nv; m++) {
              int b = d == 0 ? (m == 1 || m == 2 || m == 5 || m == 6)
                             : d == 1 ? (m == 2 || m == 3 || m == 6 || m == 7) : (m >= 4);
              double vm = b ? v1[d] : v0[d];
              acc += Nf[(size_t)m * nn + q] * (vm * vfac);
            }
            out[q] = acc;
          }
        }
        double L[3];
        for (int d = 0; d < dim; d++) L[d] = v1[d] - v0[d];
        if (elem_dxidx)
          for (int d = 0; d < dim; d++) elem_dxidx[e * dim + d] = 2.0 / L[d];
        if (elem_jac) elem_jac[e] = dim == 3 ? 0.125 * L[0] * L[1] * L[2] : 0.25 * L[0] * L[1];
      }
  // minimum distance between consecutive nodes (ggrid.ipp:977-1013)
  double tiny = 1000.0 * std::numeric_limits<double>::epsilon();
  double dmin = std::numeric_limits<double>::max();
  for (size_t q = 0; q + 1 < ndof; q++) {
    double s = 0.0;
    double dd[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) dd[d] = X[d][q + 1] - X[d][q];
    s = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
    double r = std::sqrt(s);
    if (r > tiny) dmin = std::min(dmin, r);
  }
  if (minnodedist) *minnodedist = dmin;
  const double eps = 0.125 * dmin;
  for (size_t q = 0; q < ndof; q++) {
    bool per = false, dir = false;
    bool atmax[3] = {false, false, false};
    for (int d = 0; d < dim; d++) {
      double x = X[d][q];
      bool lo = (x <= P0[d] + eps) && (x >= P0[d] - eps);
      bool hi = (x <= P1[d] + eps) && (x >= P1[d] - eps);
      atmax[d] = hi;
      int t0 = bdy[2 * d], t1 = bdy[2 * d + 1];
      if (lo) {
        if (t0 == GF_BDY_PERIODIC) per = true;
        else if (t0 != GF_BDY_NONE) dir = true;
      }
      if (hi) {
        if (t1 == GF_BDY_PERIODIC) per = true;
        else if (t1 != GF_BDY_NONE) dir = true;
      }
    }
    if (mask) mask[q] = dir ? 0.0 : 1.0;
    for (int d = 0; d < dim; d++) {
      if (xyz && xyz[d]) xyz[d][q] = X[d][q];
      if (xyzp && xyzp[d]) xyzp[d][q] = (per && atmax[d]) ? P0[d] : X[d][q];
    }
  }
  return GF_OK;
}

// GGFX local matching.  Each node owns the box [x - tol, x + tol] (per
// coordinate); node a matches node b when the boxes intersect:
// a.min <= b.max && a.max >= b.min in every coordinate (Box Intersects,
// src/tbox/spatial/bound/box.hpp:272-276).  Nodes are visited in ascending
// id; an unmapped node's matches form its group, which is assigned to every
// member (src/cdg/comm/ggfx.hpp:226-245).
//
// Candidates come from a sort of the nodes by integer cell (cells of at least
// 64*tol).  A node's matches lie within 2*tol of it, so only nodes closer than
// 3*tol to a cell face need the neighbouring cells, found by binary search.
int gf_ggfx_build(size_t ndof, int ncoord, const double *const *xyz, double tol, size_t max_dup,
                  int64_t *offs, int64_t *members, size_t *nmembers) {
  if (ncoord < 1 || ncoord > 3 || !xyz || !(tol > 0.0) || !offs || !nmembers) return GF_ERR_ARG;
  for (int d = 0; d < ncoord; d++)
    if (!xyz[d]) return GF_ERR_ARG;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < ncoord; d++) {
    lo[d] = hi[d] = ndof ? xyz[d][0] : 0.0;
    for (size_t n = 0; n < ndof; n++) {
      lo[d] = std::min(lo[d], xyz[d][n]);
      hi[d] = std::max(hi[d], xyz[d][n]);
    }
  }
  double cell = 64.0 * tol;
  for (int d = 0; d < ncoord; d++) cell = std::max(cell, (hi[d] - lo[d] + 8.0 * tol) / 1048576.0);
  auto cidx = [&](int d, double v) -> int64_t { return (int64_t)std::floor((v - lo[d] + 4.0 * tol) / cell); };
  auto pack = [](int64_t a, int64_t b, int64_t c) -> uint64_t {
    return ((uint64_t)a << 42) | ((uint64_t)b << 21) | (uint64_t)c;
  };
  std::vector<uint64_t> key(ndof);
  for (size_t n = 0; n < ndof; n++) {
    int64_t c[3] = {0, 0, 0};
    for (int d = 0; d < ncoord; d++) c[d] = cidx(d, xyz[d][n]);
    key[n] = pack(c[0], c[1], c[2]);
  }
  std::vector<int64_t> ord(ndof);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    return key[a] < key[b] || (key[a] == key[b] && a < b);
  });
  std::vector<uint64_t> skey(ndof);
  for (size_t s = 0; s < ndof; s++) skey[s] = key[ord[s]];
  std::vector<int64_t> gid(ndof, -1);  // group index per node
  std::vector<int64_t> goff(1, 0), gmem;
  gmem.reserve(ndof + ndof / 2);
  std::vector<int64_t> cand;
  auto scan_cell = [&](uint64_t k, size_t a) {
    size_t s = std::lower_bound(skey.begin(), skey.end(), k) - skey.begin();
    double amin[3], amax[3];
    for (int d = 0; d < 3; d++) {
      double v = d < ncoord ? xyz[d][a] : 0.0;
      amin[d] = v - tol;
      amax[d] = v + tol;
    }
    for (; s < ndof && skey[s] == k; s++) {
      int64_t b = ord[s];
      bool hit = true;
      for (int d = 0; d < 3 && hit; d++) {
        double v = d < ncoord ? xyz[d][b] : 0.0;
        double bmin = v - tol, bmax = v + tol;
        hit = (amin[d] <= bmax) && (amax[d] >= bmin);
      }
      if (hit) cand.push_back(b);
    }
  };
  for (size_t a = 0; a < ndof; a++) {
    if (gid[a] >= 0) continue;
    cand.clear();
    int64_t c[3] = {0, 0, 0};
    int lo_o[3] = {0, 0, 0}, hi_o[3] = {0, 0, 0};
    for (int d = 0; d < ncoord; d++) {
      double v = xyz[d][a];
      c[d] = cidx(d, v);
      if (cidx(d, v - 3.0 * tol) < c[d]) lo_o[d] = -1;
      if (cidx(d, v + 3.0 * tol) > c[d]) hi_o[d] = 1;
    }
    for (int o0 = lo_o[0]; o0 <= hi_o[0]; o0++)
      for (int o1 = lo_o[1]; o1 <= hi_o[1]; o1++)
        for (int o2 = lo_o[2]; o2 <= hi_o[2]; o2++)
          scan_cell(pack(c[0] + o0, c[1] + o1, c[2] + o2), a);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    if (max_dup && cand.size() > max_dup) return GF_ERR_ARG;
    int64_t g = (int64_t)goff.size() - 1;
    for (int64_t m : cand) {
      gid[m] = g;
      gmem.push_back(m);
    }
    goff.push_back((int64_t)gmem.size());
  }
  // Expand to per-node CSR.
  size_t tot = 0;
  for (size_t n = 0; n < ndof; n++) tot += (size_t)(goff[gid[n] + 1] - goff[gid[n]]);
  *nmembers = tot;
  if (!members) return GF_OK;
  int64_t c = 0;
  for (size_t n = 0; n < ndof; n++) {
    offs[n] = c;
    for (int64_t s = goff[gid[n]]; s < goff[gid[n] + 1]; s++) members[c++] = gmem[s];
  }
  offs[ndof] = c;
  return GF_OK;
}

}  // extern "C"
