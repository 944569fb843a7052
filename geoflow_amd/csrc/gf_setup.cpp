// gf_setup.cpp -- host-side setup for the MI355X SEM path: GLL basis, box
// grid generation and GGFX coincidence matching.  No device work here.
//
// Built with -ffp-contract=off so that node coordinates and element metrics
// round exactly as the reference's expressions do.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "geoflow_hip.h"

namespace {

typedef long double LD;

// Legendre P_n(x), P_n'(x), P_{n-1}(x) and P_{n-1}'(x) by the Jacobi
// three-term recurrence of GLLBasis::computeJacobi (gllbasis.ipp:618-667)
// with alpha = beta = 0.  The coefficients are exact integers in long double,
// and every product and quotient is the recurrence's own (a2 = 0 drops out
// exactly), so the values equal that routine's bit for bit.
struct Leg {
  LD p, dp, p1, dp1;  // P_n, P_n', P_{n-1}, P_{n-1}'
};

Leg legendre(int n, LD x) {
  Leg r{1.0L, 0.0L, 0.0L, 0.0L};
  if (n == 0) return r;
  LD pm = r.p, dm = r.dp;
  r.p = x;
  r.dp = 1.0L;
  for (int k = 2; k <= n; k++) {
    const LD fk = (LD)k;
    const LD a1 = 2.0L * fk * fk * (2.0L * fk - 2.0L);
    const LD b3 = 2.0L * fk - 2.0L;
    const LD a3 = b3 * (b3 + 1.0L) * (b3 + 2.0L);
    const LD a4 = 2.0L * (fk - 1.0L) * (fk - 1.0L) * (2.0L * fk);
    const LD ax = a3 * x;
    const LD pn = (ax * r.p - a4 * pm) / a1;
    const LD dn = (ax * r.dp - a4 * dm + a3 * r.p) / a1;
    pm = r.p;
    dm = r.dp;
    r.p = pn;
    r.dp = dn;
  }
  r.p1 = pm;
  r.dp1 = dm;
  return r;
}

// GLL nodes in long double, ascending, by GLLBasis::computeNodes'
// algorithm (gllbasis.ipp:410-501); see gf_gll_basis.
std::vector<LD> gll_nodes_ld(int p) {
  const int N = p + 1;
  std::vector<LD> x(N, 0.0L);
  // q = P_{p+1} + a P_p + b P_{p-1} vanishes at +1 and -1: a 2x2 solve with
  // the values the recurrence gives at the end points
  const Leg hi = legendre(N, 1.0L), lo = legendre(N, -1.0L);         // P_{p+1}, P_p
  const LD pm_hi = legendre(p, 1.0L).p1, pm_lo = legendre(p, -1.0L).p1;  // P_{p-1}
  const LD dt = hi.p1 * pm_lo - lo.p1 * pm_hi;
  const LD rp = -hi.p, rm = -lo.p;
  const LD ca = (rp * pm_lo - rm * pm_hi) / dt;
  const LD cb = (rm * hi.p1 - rp * lo.p1) / dt;
  // nodes from largest to smallest while they are found
  x[0] = 1.0L;
  x[p] = -1.0L;
  const int nh = (p + 1) / 2;
  // the seed angle is formed in double, as the reference's literal and
  // operands are (gllbasis.ipp:452), then used in long double
  const LD th = (LD)(3.1415926535898 / (2.0 * p + 1.0));
  const LD c2 = std::cos(2.0L * th), s2 = std::sin(2.0L * th);
  LD c1 = std::cos(th), s1 = std::sin(th);
  for (int j = 1; j <= nh; j++) {
    LD xx = c1;
    for (int it = 0; it < 128; it++) {
      const Leg a = legendre(N, xx);      // P_{p+1}, P_p and derivatives
      const Leg b = legendre(p, xx);      // P_{p-1} and its derivative in .p1/.dp1
      const LD q = a.p + ca * a.p1 + cb * b.p1;
      const LD dq = a.dp + ca * a.dp1 + cb * b.dp1;
      LD defl = 0.0L;
      for (int i = 0; i < j; i++) defl += (1.0L / (xx - x[i]));
      const LD step = -q / (dq - defl * q);
      xx += step;
      if (std::fabs(step) / std::fabs(xx) < (LD)1.0e-8) break;
    }
    x[j] = xx;
    const LD cn = c1 * c2 - s1 * s2;
    s1 = c1 * s2 + s1 * c2;
    c1 = cn;
  }
  for (int i = 0; i < nh; i++) x[N - 1 - i] = -x[i];
  if (p % 2 == 0) x[nh] = 0.0L;
  std::reverse(x.begin(), x.end());
  return x;
}

}  // namespace

extern "C" {

// GLL nodes, weights and derivative matrix with the reference's algorithm,
// so that GeoFLOW's own D and this library's agree bit for bit:
//  * nodes (GLLBasis::computeNodes, gllbasis.ipp:410-501): the interior roots
//    of q = P_{p+1} + a P_p + b P_{p-1}, with a, b chosen so q(+-1) = 0, by
//    Newton iteration with deflation of the roots already found, seeded at
//    cos((2j-1) pi/(2p+1)) (advanced by the angle-addition recurrence) and
//    stopped when the relative step falls below 1e-8 (at most 128 steps);
//    the upper half is found, mirrored, and the centre set to 0 for even p;
//  * weights 2/(p(p+1) P_p(x)^2) (computeWeights, :514-531);
//  * D(l,j) = P_p(x_l)/(P_p(x_j)(x_l-x_j) + 100 eps_ld) off the diagonal and
//    -/+ p(p+1)/4 at the two corners (computeDerivMatrix, :579-607).
// All of it in long double, cast to double at the end, except where the
// reference's own expressions are double: the seed angle (its literal
// 3.1415926535898 divided by 2p+1), the stopping threshold and the weight
// factor 2/(p(p+1)).
int gf_gll_basis(int p, double *xi, double *w, double *D) {
  if (p < 1 || p > 64 || !xi || !w || !D) return GF_ERR_ARG;
  const int N = p + 1;
  const std::vector<LD> x = gll_nodes_ld(p);
  std::vector<LD> P(N);
  const LD fact = (LD)(2.0 / (p * (p + 1.0)));  // formed in double (gllbasis.ipp:520)
  for (int i = 0; i < N; i++) {
    P[i] = legendre(p, x[i]).p;
    xi[i] = (double)x[i];
    w[i] = (double)(P[i] == 0.0L ? 0.0L : fact / (P[i] * P[i]));
  }
  const LD ttiny = 100.0L * std::numeric_limits<LD>::epsilon();
  for (int j = 0; j < N; j++)
    for (int l = 0; l < N; l++) {
      LD v = 0.0L;
      if (l != j) v = P[l] / (P[j] * (x[l] - x[j]) + ttiny);
      D[l + (size_t)j * N] = (double)v;
    }
  const LD d00 = -0.25L * p * (p + 1.0L);
  D[0] = (double)d00;
  D[(size_t)p + (size_t)p * N] = (double)(-d00);
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

// ---------------------------------------------------------------------------
// Icosahedral sphere grid (GGridIcos, 2-D embedded quads on a sphere)
// ---------------------------------------------------------------------------
namespace {

const double kPi = 3.14159265358979323846264338328;  // PI (src/cdg/include/gtypes.h:92)

struct V3 {
  double x, y, z;
};
inline V3 vadd(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vsub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vmul(const V3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
// GTPoint::norm: sqrt(((0 + x^2) + y^2) + z^2)
inline double vnorm(const V3 &a) {
  double s = 0.0;
  s += a.x * a.x;
  s += a.y * a.y;
  s += a.z * a.z;
  return std::sqrt(s);
}

// Radial projection through latitude/longitude (project2sphere,
// ggrid_icos.ipp:1457-1556); `r` is the norm as the caller computes it.
inline V3 to_sphere(double x, double y, double z, double r, double rad) {
  const double lat = std::asin(z / r);
  double lon = std::atan2(y, x);
  lon = lon < 0.0 ? 2.0 * kPi + lon : lon;
  return {rad * std::cos(lat) * std::cos(lon), rad * std::cos(lat) * std::sin(lon),
          rad * std::sin(lat)};
}

// GTVector::sortincreasing (gtvector.ipp:2965-2999, 3013-3243): quicksort with
// the middle element as pivot that sorts `a` IN PLACE and permutes `is`
// alongside (is starts as 0..n-1).
void qsort_s2l(double *a, size_t *is, long lo, long hi) {
  if (lo >= hi) return;
  const long pi = lo + (hi - lo) / 2;
  const double pv = a[pi];
  std::swap(a[pi], a[lo]);
  std::swap(is[pi], is[lo]);
  long i = lo + 1, j = hi;
  while (i <= j) {
    while (i <= j && a[i] <= pv) i++;
    while (i <= j && a[j] > pv) j--;
    if (i < j) {
      std::swap(a[i], a[j]);
      std::swap(is[i], is[j]);
    }
  }
  std::swap(a[i - 1], a[lo]);
  std::swap(is[i - 1], is[lo]);
  qsort_s2l(a, is, lo, i - 2);
  qsort_s2l(a, is, i, hi);
}
void sort_increasing(double *a, size_t *is, size_t n) {
  for (size_t k = 0; k < n; k++) is[k] = k;
  qsort_s2l(a, is, 0, (long)n - 1);
}

// xyz2spherical (ggrid_icos.ipp:1698-1720): latitude and longitude in [0,2pi).
inline void latlon(const V3 &v, double &lat, double &lon) {
  const double r = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  lat = std::asin(v.z / r);
  lon = std::atan2(v.y, v.x);
  lon = lon < 0.0 ? 2.0 * kPi + lon : lon;
}

// GLL basis data in long double for basis/derivative evaluation at nodes.
struct LDBasis {
  int p;
  std::vector<LD> x, Pn;
  std::vector<double> xd;
};

// P_N(x) and P_N'(x) by GLLBasis::computeJacobi's recurrence (see legendre).
void jacobi_ld(int N, LD x, LD &poly, LD &pder) {
  const Leg r = legendre(N, x);
  poly = r.p;
  pder = r.dp;
}

// h_i(eta) of GLLBasis::evalBasis (2) (gllbasis.ipp:1300-1327).
double eval_basis(const LDBasis &b, int i, double eta) {
  const int Np = b.p;
  const LD fact = -1.0 / (Np * (Np + 1.0));  // formed in double, as the reference does
  const LD ttiny = 100.0L * std::numeric_limits<LD>::epsilon();
  const LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  const LD xi = (LD)eta;
  double f = 1.0;
  if (xi < -1.0L || xi > 1.0L) f = 0.0;
  else if (std::fabs(xi - b.x[i]) > tetiny) {
    LD P, dP;
    jacobi_ld(Np, xi, P, dP);
    const LD g = (1.0L - xi * xi) / (xi - b.x[i] + ttiny);
    f = (double)(fact * g * dP / (b.Pn[i] + ttiny));
  }
  return f;
}

// h_j(eta) of GLLBasis::evalBasis (3), the matrix form used by
// init_qdealias (gllbasis.ipp:1341-1377): no ttiny shifts.
double eval_basis_m(const LDBasis &b, int j, double eta) {
  const int Np = b.p;
  const LD fact = -1.0 / (Np * (Np + 1.0));
  const LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  const LD xi = (LD)eta;
  double f = 1.0;
  if (xi < -1.0L || xi > 1.0L) f = 0.0;
  else if (std::fabs(xi - b.x[j]) > tetiny) {
    LD P, dP;
    jacobi_ld(Np, xi, P, dP);
    const LD g = (1.0L - xi * xi) / (xi - b.x[j]);
    f = (double)(fact * g * dP / b.Pn[j]);
  }
  return f;
}

// h_j'(eta) of GLLBasis::evalDBasis (4) (gllbasis.ipp:1628-1665).
double eval_dbasis(const LDBasis &b, int j, double eta) {
  const int Np = b.p;
  const LD fact = -1.0 / (Np * (Np + 1.0));  // formed in double, as the reference does
  const LD tetiny = 100.0L * std::numeric_limits<double>::epsilon();
  const LD xi = (LD)eta;
  const LD g1 = xi - b.x[j];
  double f = 0.0;
  if (xi == -1.0L && j == 0) f = (double)(-0.25L * Np * (Np + 1.0L));
  else if (xi == 1.0L && j == Np) f = (double)(0.25L * Np * (Np + 1.0L));
  else if (std::fabs(g1) > tetiny) {
    LD P, dP;
    jacobi_ld(Np, xi, P, dP);
    const LD gfact = fact / b.Pn[j];
    const LD g1i = 1.0L / g1;
    const LD g2 = (1.0L - xi * xi) * g1i;
    const LD pdd = 2.0L * xi * dP - Np * (Np + 1.0L) * P;
    f = (double)(gfact * g1i * (pdd - (2.0L * xi + g2) * dP));
  }
  return f;
}

LDBasis make_ldbasis(int p) {
  LDBasis b;
  b.p = p;
  b.xd.resize(p + 1);
  std::vector<double> w(p + 1), D((size_t)(p + 1) * (p + 1));
  gf_gll_basis(p, b.xd.data(), w.data(), D.data());
  b.x.resize(p + 1);
  b.Pn.resize(p + 1);
  // the long-double nodes themselves (GLLBasis::xiNodes_), as evalBasis and
  // evalDBasis use them
  b.x = gll_nodes_ld(p);
  b.Pn.resize(p + 1);
  for (int i = 0; i <= p; i++) {
    LD dP;
    jacobi_ld(p, b.x[i], b.Pn[i], dP);
  }
  return b;
}

}  // namespace

extern "C" {

// GGridIcos 2-D for one rank (src/cdg/grid/ggrid_icos.ipp): the base
// icosahedron (init2d, :139-230), 'Lagrangian' refinement of every face into
// (nrows+1)^2 triangles (lagrefine/lagvert/interleave, :254-328, 2127-2173),
// projection to the unit sphere, triangle vertex ordering (order_triangles,
// :2202-2258, including its in-place sort), then three quads per triangle
// (do_elems2d, :379-513) whose vertices are projected and ordered in gnomonic
// coordinates (reorderverts2d/cart2gnomonic, :1733-1990), nodes from the
// bilinear shape functions projected to radius R, and the embedded metrics
// of GElem_base::dogeom2d/Jac/inv (src/cdg/sem/gelem_base.cpp:877-955,
// 1066-1170, 1249-1300): rij = dx^k/dxi^m (the third column is the position,
// from the embedded shape functions, gshapefcn_embed.ipp:356-390),
// Jac = |det rij| / |x|, dXi/dX = rij^-1.
//   refine: 0 bisection (nrows = 2^ilevel - 1), 1 Lagrangian (nrows = ilevel)
// Outputs (caller-allocated, ndof = nelems (p+1)^2; may be NULL to query):
//   nelems, xyz[3], dxidx[6] (dXi_i/dX_j at [i*3+j], i < 2), jac, minnodedist
int gf_icos_grid(int p, int ilevel, int refine, double radius, size_t *nelems,
                 double *const *xyz, double *const *dxidx, double *jac, double *minnodedist) {
  if (p < 1 || p > 64 || ilevel < 0 || ilevel > 12 || (refine != 0 && refine != 1) ||
      !(radius > 0.0) || !nelems)
    return GF_ERR_ARG;
  const long nrows = refine == 0 ? (1L << ilevel) - 1 : ilevel;
  const size_t ntri = 20 * (size_t)(nrows * (nrows + 2) + 1);
  *nelems = 3 * ntri;
  if (!xyz) return GF_OK;
  for (int k = 0; k < 3; k++)
    if (!xyz[k]) return GF_ERR_ARG;

  // base icosahedron, vertex at the top (init2d, ggrid_icos.ipp:163-203)
  static const double fv0[12][3] = {
      {0.000000000000000, 0.000000000000000, 1.000000000000000},
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
  static const int ifv0[20][3] = {{0, 1, 8},  {0, 8, 4},  {0, 4, 5},  {0, 5, 9},  {0, 9, 1},
                                  {1, 6, 8},  {8, 6, 10}, {8, 10, 4}, {4, 10, 2}, {4, 2, 5},
                                  {5, 2, 11}, {5, 11, 9}, {9, 11, 7}, {9, 7, 1},  {1, 7, 6},
                                  {3, 6, 7},  {3, 7, 11}, {3, 11, 2}, {3, 2, 10}, {3, 10, 6}};
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  const double vfact = 1.0 / std::sqrt(phi * phi + 1.0);
  V3 bv[12];
  for (int i = 0; i < 12; i++) bv[i] = {fv0[i][0] * vfact, fv0[i][1] * vfact, fv0[i][2] * vfact};

  // refined triangles (lagrefine): rows of 'Lagrangian' vertices
  std::vector<std::array<V3, 3>> tri(ntri);
  {
    size_t n = 0;
    const double rfact = 1.0 / (double)(nrows + 1);
    auto row = [&](const V3 &a, const V3 &b, const V3 &c, long I, std::vector<V3> &R) {
      R.resize(I + 1);
      const double xI = (double)I;
      const V3 rL = vadd(a, vmul(vsub(b, a), xI * rfact));
      const V3 rR = vadd(a, vmul(vsub(c, a), xI * rfact));
      const double f = I > 0 ? 1.0 / (double)I : 1.0;
      for (long j = 0; j <= I; j++) R[j] = vadd(rL, vmul(vsub(rR, rL), (double)j * f));
    };
    std::vector<V3> R0, R1, Rz;
    for (int t = 0; t < 20; t++) {
      const V3 a = bv[ifv0[t][0]], b = bv[ifv0[t][1]], c = bv[ifv0[t][2]];
      for (long l = 0; l < nrows + 1; l++) {
        row(a, b, c, l, R0);
        row(a, b, c, l + 1, R1);
        Rz.resize(2 * l + 3);
        for (long j = 0; j < 2 * l + 3; j++) Rz[j] = (j % 2 == 0) ? R1[j / 2] : R0[(j - 1) / 2];
        for (long m = 0; m < 2 * l + 1; m++) tri[n++] = {Rz[m], Rz[m + 1], Rz[m + 2]};
      }
    }
  }
  // project to the unit sphere, then order the vertices (order_triangles)
  std::vector<int> iup(ntri, 0);
  for (size_t t = 0; t < ntri; t++) {
    for (int j = 0; j < 3; j++) {
      const V3 v = tri[t][j];
      tri[t][j] = to_sphere(v.x, v.y, v.z, vnorm(v), 1.0);
    }
    const std::array<V3, 3> cv = tri[t];
    double lat[3], lon[3];
    for (int j = 0; j < 3; j++) latlon(cv[j], lat[j], lon[j]);
    size_t is[3];
    sort_increasing(lon, is, 3);  // lon is now sorted; is refers to the input order
    if (std::fabs(lon[is[0]] - lon[is[2]]) < kPi)
      for (int j = 0; j < 3; j++)
        if (lon[j] > 1.5 * kPi && lon[j] <= 2.0 * kPi) lon[j] -= 2.0 * kPi;
    sort_increasing(lon, is, 3);
    if (lat[is[1]] > lat[is[0]] && lat[is[1]] > lat[is[2]]) {
      tri[t] = {cv[is[0]], cv[is[2]], cv[is[1]]};
      iup[t] = 1;
    }
    if (lat[is[1]] < lat[is[0]] && lat[is[1]] < lat[is[2]]) tri[t] = {cv[is[1]], cv[is[2]], cv[is[0]]};
  }

  // elements: 3 quads per triangle
  const int N = p + 1;
  const size_t nn = (size_t)N * N;
  const LDBasis lb = make_ldbasis(p);
  const std::vector<double> &xi = lb.xd;
  // bilinear shape functions N_m at the nodes (GShapeFcn_linear::Ni_2d)
  std::vector<double> Nf(4 * nn);
  for (int m = 0; m < 4; m++)
    for (int j = 0, q = 0; j < N; j++)
      for (int i = 0; i < N; i++, q++) {
        const double rm = 1.0 - xi[i], rp = 1.0 + xi[i], sm = 1.0 - xi[j], sp = 1.0 + xi[j];
        Nf[m * nn + q] = m == 0 ? rm * sm : (m == 1 ? rp * sm : (m == 2 ? rp * sp : rm * sp));
      }
  // embedded shape-function derivatives at the nodes (GShapeFcn_embed::dNdXi_2d)
  std::vector<double> hb((size_t)N * N), hd((size_t)N * N);  // [basis i][node]
  for (int i = 0; i < N; i++)
    for (int q = 0; q < N; q++) {
      hb[(size_t)i * N + q] = eval_basis(lb, i, xi[q]);
      hd[(size_t)i * N + q] = eval_dbasis(lb, i, xi[q]);
    }
  const double third = 1.0 / 3.0;
  const double geps = 10.0 * std::numeric_limits<double>::epsilon();
  size_t e = 0;
  std::vector<double> X[3];
  for (int k = 0; k < 3; k++) X[k].resize(nn);
  for (size_t t = 0; t < ntri; t++) {
    const V3 v1 = tri[t][0], v2 = tri[t][1], v3 = tri[t][2];
    const V3 ct = vmul(vadd(v1, vadd(v2, v3)), third);
    const V3 m12 = vmul(vadd(v1, v2), 0.5), m13 = vmul(vadd(v1, v3), 0.5),
             m23 = vmul(vadd(v2, v3), 0.5);
    for (int j = 0; j < 3; j++, e++) {
      V3 cv[4];
      if (iup[t] == 1) {
        if (j == 0) { cv[0] = v1; cv[1] = m12; cv[2] = ct; cv[3] = m13; }
        else if (j == 1) { cv[0] = m12; cv[1] = v2; cv[2] = m23; cv[3] = ct; }
        else { cv[0] = ct; cv[1] = m23; cv[2] = v3; cv[3] = m13; }
      } else {
        if (j == 0) { cv[0] = v1; cv[1] = m12; cv[2] = ct; cv[3] = m13; }
        else if (j == 1) { cv[0] = ct; cv[1] = m12; cv[2] = v2; cv[3] = m23; }
        else { cv[0] = m13; cv[1] = ct; cv[2] = m23; cv[3] = v3; }
      }
      for (int m = 0; m < 4; m++) cv[m] = to_sphere(cv[m].x, cv[m].y, cv[m].z, vnorm(cv[m]), 1.0);
      // reorderverts2d: order the quad's vertices in gnomonic coordinates
      const V3 c = vmul(vadd(vadd(vadd(cv[0], cv[1]), cv[2]), cv[3]), 0.25);
      const double latc = std::asin(c.z);
      double lonc = std::atan2(c.y, c.x);
      lonc = lonc < 0.0 ? 2 * kPi + lonc : lonc;
      double gx[4], gy[4];
      for (int m = 0; m < 4; m++) {
        const double r = vnorm(cv[m]);
        const double la = std::asin(cv[m].z / r);
        double lo = std::atan2(cv[m].y, cv[m].x);
        lo = lo < 0.0 ? 2.0 * kPi + lo : lo;
        double den = std::sin(latc) * std::sin(la) + std::cos(latc) * std::cos(la) * std::cos(lo - lonc);
        den = std::fabs(den) < geps ? 0.0 : 1.0 / den;
        gx[m] = 1.0 * std::cos(la) * std::sin(lo - lonc) * den;
        gy[m] = 1.0 * (std::cos(latc) * std::sin(la) - std::sin(latc) * std::cos(la) * std::cos(lo - lonc)) * den;
      }
      double sx[4] = {gx[0], gx[1], gx[2], gx[3]};
      size_t I[4], ord[4];
      sort_increasing(sx, I, 4);
      if (gy[I[0]] < gy[I[1]]) { ord[0] = I[0]; ord[3] = I[1]; }
      else { ord[0] = I[1]; ord[3] = I[0]; }
      if (gy[I[2]] < gy[I[3]]) { ord[1] = I[2]; ord[2] = I[3]; }
      else { ord[1] = I[3]; ord[2] = I[2]; }
      V3 gv[4];
      for (int m = 0; m < 4; m++) gv[m] = cv[ord[m]];
      // nodes: sum_m N_m (v_m / 4), projected to radius R
      const size_t base = e * nn;
      for (size_t q = 0; q < nn; q++) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < 4; m++) {
          acc[0] += Nf[m * nn + q] * (gv[m].x * 0.25);
          acc[1] += Nf[m * nn + q] * (gv[m].y * 0.25);
          acc[2] += Nf[m * nn + q] * (gv[m].z * 0.25);
        }
        const double r = std::sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
        const V3 s = to_sphere(acc[0], acc[1], acc[2], r, radius);
        X[0][q] = s.x;
        X[1][q] = s.y;
        X[2][q] = s.z;
        xyz[0][base + q] = s.x;
        xyz[1][base + q] = s.y;
        xyz[2][base + q] = s.z;
      }
      if (!dxidx && !jac) continue;
      // rij(k,m) = sum_I (h_m-derivative product) x^k_I, I in (j outer, i inner)
      for (int qj = 0, q = 0; qj < N; qj++)
        for (int qi = 0; qi < N; qi++, q++) {
          double r[3][3];
          for (int k = 0; k < 3; k++)
            for (int m = 0; m < 3; m++) {
              double tmp = 0.0;
              for (int bj = 0, n = 0; bj < N; bj++)
                for (int bi = 0; bi < N; bi++, n++) {
                  const double d0 = m == 0 ? hd[(size_t)bi * N + qi] : hb[(size_t)bi * N + qi];
                  const double d1 = m == 1 ? hd[(size_t)bj * N + qj] : hb[(size_t)bj * N + qj];
                  double dn = d0 * d1;
                  dn *= X[k][n];
                  tmp += dn;
                }
              r[k][m] = tmp;
            }
          const double det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
                             r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
                             r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
          if (jac) {
            const double x = r[0][2], y = r[1][2], z = r[2][2];
            jac[base + q] = std::fabs(det) / std::sqrt(x * x + y * y + z * z);
          }
          if (dxidx) {
            const double ij = 1.0 / det;
            const double i00 = (r[1][1] * r[2][2] - r[1][2] * r[2][1]) * ij;
            const double i01 = -(r[0][1] * r[2][2] - r[2][1] * r[0][2]) * ij;
            const double i02 = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) * ij;
            const double i10 = -(r[1][0] * r[2][2] - r[2][0] * r[1][2]) * ij;
            const double i11 = (r[0][0] * r[2][2] - r[2][0] * r[0][2]) * ij;
            const double i12 = -(r[0][0] * r[1][2] - r[0][2] * r[1][0]) * ij;
            const double iv[6] = {i00, i01, i02, i10, i11, i12};
            for (int c2 = 0; c2 < 6; c2++)
              if (dxidx[c2]) dxidx[c2][base + q] = iv[c2];
          }
        }
    }
  }
  if (minnodedist) {  // GGrid::find_min_dist (ggrid.ipp:977-1013)
    const size_t ndof = *nelems * nn;
    const double tiny = 1000.0 * std::numeric_limits<double>::epsilon();
    double dmin = std::numeric_limits<double>::max();
    for (size_t q = 0; q + 1 < ndof; q++) {
      const double dx = xyz[0][q + 1] - xyz[0][q], dy = xyz[1][q + 1] - xyz[1][q],
                   dz = xyz[2][q + 1] - xyz[2][q];
      const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (rr > tiny) dmin = std::min(dmin, rr);
    }
    *minnodedist = dmin;
  }
  return GF_OK;
}

// GE_DEFORMED geometry of 2-D elements from their node coordinates:
// GElem_base::dogeom2d (src/cdg/sem/gelem_base.cpp:877-955): R(k,m) = dx_k/dxi_m
// from the tensor-product GLL shape-function derivatives at the nodes
// (GLLBasis::evalBasis / evalDBasis), jac = R00 R11 - R01 R10 (GElem_base::Jac,
// 2x2 branch without absolute value, gelem_base.cpp:1121-1136) and dXi/dX
// from GElem_base::inv (gelem_base.cpp:1249-1266).  That 2x2 inverse returns
// the transpose of R^-1 (its (0,1) entry is -R10/jac); reference_inverse != 0
// keeps it, so the metrics equal the reference's, 0 gives the true inverse.
//   xyz[2]: nelems * (p+1)^2 node coordinates; dxidx[4] at [i*2+j] =
//   dXi_i/dX_j per node; jac per node.
int gf_deformed_geometry2d(int p, size_t nelems, const double *const *xyz, int reference_inverse,
                           double *const *dxidx, double *jac) {
  if (p < 1 || !xyz || !xyz[0] || !xyz[1] || !dxidx || !jac) return GF_ERR_ARG;
  for (int k = 0; k < 4; k++)
    if (!dxidx[k]) return GF_ERR_ARG;
  const int N = p + 1;
  const LDBasis lb = make_ldbasis(p);
  // tensor-product shape-function values and derivatives at the nodes,
  // [basis i][node q]
  std::vector<double> hb((size_t)N * N), hd((size_t)N * N);
  for (int i = 0; i < N; i++)
    for (int q = 0; q < N; q++) {
      hb[(size_t)i * N + q] = eval_basis(lb, i, lb.xd[q]);
      hd[(size_t)i * N + q] = eval_dbasis(lb, i, lb.xd[q]);
    }
  const size_t nn = (size_t)N * N;
  for (size_t e = 0; e < nelems; e++) {
    const size_t base = e * nn;
    for (int qj = 0, q = 0; qj < N; qj++)
      for (int qi = 0; qi < N; qi++, q++) {
        // R(k,m) = sum_I dN_I/dxi_m x_k[I], I in (j outer, i inner)
        double r[2][2];
        for (int k = 0; k < 2; k++)
          for (int m = 0; m < 2; m++) {
            double tmp = 0.0;
            for (int bj = 0, n = 0; bj < N; bj++)
              for (int bi = 0; bi < N; bi++, n++) {
                const double d0 = m == 0 ? hd[(size_t)bi * N + qi] : hb[(size_t)bi * N + qi];
                const double d1 = m == 1 ? hd[(size_t)bj * N + qj] : hb[(size_t)bj * N + qj];
                double dn = d0 * d1;
                dn *= xyz[k][base + n];
                tmp += dn;
              }
            r[k][m] = tmp;
          }
        const size_t id = base + q;
        const double jv = r[0][0] * r[1][1] - r[0][1] * r[1][0];
        const double ij = 1.0 / jv;
        jac[id] = jv;
        dxidx[0][id] = r[1][1] * ij;
        dxidx[3][id] = r[0][0] * ij;
        dxidx[1][id] = (reference_inverse ? -r[1][0] : -r[0][1]) * ij;
        dxidx[2][id] = (reference_inverse ? -r[0][1] : -r[1][0]) * ij;
      }
  }
  return GF_OK;
}

// GGrid::find_min_dist over consecutive local nodes (ggrid.ipp:977-1013), 2-D.
static double min_dist2d(const double *const *xyz, size_t ndof) {
  const double tiny = 1000.0 * std::numeric_limits<double>::epsilon();
  double dmin = std::numeric_limits<double>::max();
  for (size_t q = 0; q + 1 < ndof; q++) {
    const double dx = xyz[0][q + 1] - xyz[0][q], dy = xyz[1][q + 1] - xyz[1][q];
    const double rr = std::sqrt(dx * dx + dy * dy);
    if (rr > tiny) dmin = std::min(dmin, rr);
  }
  return dmin;
}

// GIO state files (src/cdg/io/gio.ipp): the header of GIO::read_header
// (gio.ipp:588-640) -- int ivers, int dim, size_t nelems, int porder[rows*dim]
// (rows = 1 for ivers 0, else nelems), int gtype, size_t cycle, double time,
// int multivar, size_t keys[nelems] -- of GIO::sz_header bytes
// (gio.ipp:678-690), followed by the fields' doubles (one field, or all of
// them in order when multivar is set).  Native byte order, as fwrite.
int gf_gio_write(const char *path, const gf_gio_info *info, const int *porder,
                 const uint64_t *keys, const double *const *fields, int nfields, size_t npts) {
  if (!path || !info || !porder || !fields || nfields < 1 || info->dim < 1 || info->dim > 3)
    return GF_ERR_ARG;
  if (!info->multivar && nfields != 1) return GF_ERR_ARG;
  for (int f = 0; f < nfields; f++)
    if (!fields[f]) return GF_ERR_ARG;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return GF_ERR_ARG;
  const int ivers = info->ivers, dim = info->dim, gtype = info->gtype, multi = info->multivar;
  const size_t nelems = (size_t)info->nelems, cycle = (size_t)info->cycle;
  const double time = info->time;
  const size_t nrow = ivers == 0 ? 1 : nelems;
  bool ok = std::fwrite(&ivers, sizeof(int), 1, fp) == 1 && std::fwrite(&dim, sizeof(int), 1, fp) == 1 &&
            std::fwrite(&nelems, sizeof(size_t), 1, fp) == 1 &&
            std::fwrite(porder, sizeof(int), nrow * dim, fp) == nrow * dim &&
            std::fwrite(&gtype, sizeof(int), 1, fp) == 1 &&
            std::fwrite(&cycle, sizeof(size_t), 1, fp) == 1 &&
            std::fwrite(&time, sizeof(double), 1, fp) == 1 &&
            std::fwrite(&multi, sizeof(int), 1, fp) == 1;
  for (size_t e = 0; ok && e < nelems; e++) {
    const size_t k = keys ? (size_t)keys[e] : e;
    ok = std::fwrite(&k, sizeof(size_t), 1, fp) == 1;
  }
  for (int f = 0; ok && f < nfields; f++) ok = std::fwrite(fields[f], sizeof(double), npts, fp) == npts;
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? GF_OK : GF_ERR_ARG;
}

// Reads a GIO file written as above.  Call with fields == NULL to read the
// header only (porder_cap / keys_cap bound the arrays, which may be NULL);
// then with nfields field buffers of npts doubles.
int gf_gio_read(const char *path, gf_gio_info *info, int *porder, size_t porder_cap,
                uint64_t *keys, size_t keys_cap, double *const *fields, int nfields, size_t npts) {
  if (!path || !info) return GF_ERR_ARG;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return GF_ERR_ARG;
  int ivers = 0, dim = 0, gtype = 0, multi = 0;
  size_t nelems = 0, cycle = 0;
  double time = 0.0;
  bool ok = std::fread(&ivers, sizeof(int), 1, fp) == 1 && std::fread(&dim, sizeof(int), 1, fp) == 1 &&
            std::fread(&nelems, sizeof(size_t), 1, fp) == 1 && dim >= 1 && dim <= 3;
  const size_t nrow = ok ? (ivers == 0 ? 1 : nelems) : 0;
  std::vector<int> po(ok ? nrow * dim : 0);
  ok = ok && std::fread(po.data(), sizeof(int), po.size(), fp) == po.size() &&
       std::fread(&gtype, sizeof(int), 1, fp) == 1 && std::fread(&cycle, sizeof(size_t), 1, fp) == 1 &&
       std::fread(&time, sizeof(double), 1, fp) == 1 && std::fread(&multi, sizeof(int), 1, fp) == 1;
  std::vector<size_t> ks(ok ? nelems : 0);
  ok = ok && std::fread(ks.data(), sizeof(size_t), ks.size(), fp) == ks.size();
  if (ok) {
    info->ivers = ivers;
    info->dim = dim;
    info->nelems = nelems;
    info->gtype = gtype;
    info->cycle = cycle;
    info->time = time;
    info->multivar = multi;
    info->nporder = po.size();
    if (porder) std::copy(po.begin(), po.begin() + std::min(porder_cap, po.size()), porder);
    if (keys)
      for (size_t e = 0; e < std::min(keys_cap, ks.size()); e++) keys[e] = (uint64_t)ks[e];
  }
  for (int f = 0; ok && fields && f < nfields; f++)
    ok = fields[f] && std::fread(fields[f], sizeof(double), npts, fp) == npts;
  std::fclose(fp);
  return ok ? GF_OK : GF_ERR_ARG;
}

// GE_DEFORMED geometry of 3-D elements from their node coordinates:
// GElem_base::dogeom3d (src/cdg/sem/gelem_base.cpp:972-1048): R(l,m) =
// dx_l/dxi_m from the tensor-product GLL shape-function derivatives at the
// nodes, jac = |det R| (GElem_base::Jac, 3x3 branch, gelem_base.cpp:1148-1168)
// and dXi/dX = R^-1 by cofactors over det R (GElem_base::inv,
// gelem_base.cpp:1268-1283; the 3x3 branch is the true inverse).
//   xyz[3]: nelems * (p+1)^3 coordinates; dxidx[9] at [i*3+j] = dXi_i/dX_j.
int gf_deformed_geometry3d(int p, size_t nelems, const double *const *xyz,
                           double *const *dxidx, double *jac) {
  if (p < 1 || !xyz || !xyz[0] || !xyz[1] || !xyz[2] || !dxidx || !jac) return GF_ERR_ARG;
  for (int k = 0; k < 9; k++)
    if (!dxidx[k]) return GF_ERR_ARG;
  const int N = p + 1;
  const LDBasis lb = make_ldbasis(p);
  std::vector<double> hb((size_t)N * N), hd((size_t)N * N);  // [basis i][node q]
  for (int i = 0; i < N; i++)
    for (int q = 0; q < N; q++) {
      hb[(size_t)i * N + q] = eval_basis(lb, i, lb.xd[q]);
      hd[(size_t)i * N + q] = eval_dbasis(lb, i, lb.xd[q]);
    }
  const size_t nn = (size_t)N * N * N;
  for (size_t e = 0; e < nelems; e++) {
    const size_t base = e * nn;
    for (int qk = 0, q = 0; qk < N; qk++)
      for (int qj = 0; qj < N; qj++)
        for (int qi = 0; qi < N; qi++, q++) {
          double r[3][3];
          for (int m = 0; m < 3; m++)
            for (int l = 0; l < 3; l++) {
              // R(l,m) = sum_I dN_I/dxi_m x_l[I], I in (k, j, i) order
              double tmp = 0.0;
              for (int bk = 0, n = 0; bk < N; bk++)
                for (int bj = 0; bj < N; bj++)
                  for (int bi = 0; bi < N; bi++, n++) {
                    const double d0 = m == 0 ? hd[(size_t)bi * N + qi] : hb[(size_t)bi * N + qi];
                    const double d1 = m == 1 ? hd[(size_t)bj * N + qj] : hb[(size_t)bj * N + qj];
                    const double d2 = m == 2 ? hd[(size_t)bk * N + qk] : hb[(size_t)bk * N + qk];
                    double dn = d0 * d1 * d2;
                    dn *= xyz[l][base + n];
                    tmp += dn;
                  }
              r[l][m] = tmp;
            }
          auto G = [&](int a, int b) { return r[a][b]; };
          const double det = G(0, 0) * (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) -
                             G(0, 1) * (G(1, 0) * G(2, 2) - G(1, 2) * G(2, 0)) +
                             G(0, 2) * (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0));
          const size_t id = base + q;
          jac[id] = std::fabs(det);
          const double ij = 1.0 / det;
          dxidx[0][id] = (G(1, 1) * G(2, 2) - G(1, 2) * G(2, 1)) * ij;
          dxidx[1][id] = -(G(0, 1) * G(2, 2) - G(2, 1) * G(0, 2)) * ij;
          dxidx[2][id] = (G(0, 1) * G(1, 2) - G(0, 2) * G(1, 1)) * ij;
          dxidx[3][id] = -(G(1, 0) * G(2, 2) - G(2, 0) * G(1, 2)) * ij;
          dxidx[4][id] = (G(0, 0) * G(2, 2) - G(2, 0) * G(0, 2)) * ij;
          dxidx[5][id] = -(G(0, 0) * G(1, 2) - G(0, 2) * G(1, 0)) * ij;
          dxidx[6][id] = (G(1, 0) * G(2, 1) - G(1, 1) * G(2, 0)) * ij;
          dxidx[7][id] = -(G(0, 0) * G(2, 1) - G(0, 1) * G(2, 0)) * ij;
          dxidx[8][id] = (G(0, 0) * G(1, 1) - G(1, 0) * G(0, 1)) * ij;
        }
  }
  return GF_OK;
}

// 3-D box of GE_DEFORMED hexahedra whose bottom (z = P0z) follows
//   h(x,y) = amp (1 - cos(2 pi (x-P0x)/Lx)) (1 - cos(2 pi (y-P0y)/Ly)) / 4,
// every node moved by z' = z + h(x,y) (P1z - z)/(P1z - P0z); arguments as
// gf_terrain_box_grid with 3 directions (z may not be periodic) and
// dxidx[9] at [i*3+j]; geometry as gf_deformed_geometry3d.
int gf_terrain_box_grid3d(int p, const int *ne, const double *P0, const double *P1,
                          const int *bdy, double amp, double *const *xyz, double *const *xyzp,
                          double *mask, double *const *dxidx, double *jac, double *minnodedist) {
  if (p < 1 || !ne || !P0 || !P1 || !bdy || !xyz || !xyz[0] || !xyz[1] || !xyz[2] || !dxidx ||
      !jac)
    return GF_ERR_ARG;
  if (bdy[4] == GF_BDY_PERIODIC || bdy[5] == GF_BDY_PERIODIC) return GF_ERR_ARG;
  const int pp[3] = {p, p, p};
  int rc = gf_box_grid(3, pp, ne, P0, P1, bdy, xyz, xyzp, mask, nullptr, nullptr, nullptr);
  if (rc) return rc;
  const int N = p + 1;
  const double Lx = P1[0] - P0[0], Ly = P1[1] - P0[1], H = P1[2] - P0[2];
  const size_t nel = (size_t)ne[0] * ne[1] * ne[2], ndof = nel * N * N * N;
  auto move = [&](double *const *c) {
    for (size_t q = 0; q < ndof; q++) {
      const double h = amp * 0.25 * (1.0 - std::cos(2.0 * kPi * (c[0][q] - P0[0]) / Lx)) *
                       (1.0 - std::cos(2.0 * kPi * (c[1][q] - P0[1]) / Ly));
      c[2][q] = c[2][q] + h * (P1[2] - c[2][q]) / H;
    }
  };
  move(xyz);
  if (xyzp && xyzp[0] && xyzp[1] && xyzp[2]) move(xyzp);
  rc = gf_deformed_geometry3d(p, nel, xyz, dxidx, jac);
  if (rc) return rc;
  if (minnodedist) {  // GGrid::find_min_dist (ggrid.ipp:977-1013)
    const double tiny = 1000.0 * std::numeric_limits<double>::epsilon();
    double dmin = std::numeric_limits<double>::max();
    for (size_t q = 0; q + 1 < ndof; q++) {
      const double dx = xyz[0][q + 1] - xyz[0][q], dy = xyz[1][q + 1] - xyz[1][q],
                   dz = xyz[2][q + 1] - xyz[2][q];
      const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (rr > tiny) dmin = std::min(dmin, rr);
    }
    *minnodedist = dmin;
  }
  return GF_OK;
}

// 2-D box of GE_DEFORMED quads following a terrain profile: the nodes of the
// regular box (gf_box_grid) are moved by the terrain-following map
//   y' = y + h(x) (P1y - y) / (P1y - P0y),  h(x) = amp (1 - cos(2 pi (x - P0x) / Lx)) / 2,
// which raises the bottom boundary to h(x) and keeps the top; geometry as
// gf_deformed_geometry2d.
//   p, ne[2], P0[2], P1[2], bdy[4] as gf_box_grid (y may not be periodic);
//   xyz[2], xyzp[2] (periodized), mask: as gf_box_grid, for the moved nodes;
//   dxidx[4] at [i*2+j] = dXi_i/dX_j per node; jac per node; minnodedist of
//   the moved nodes (GGrid::find_min_dist, src/cdg/grid/ggrid.ipp:977-1013).
int gf_terrain_box_grid(int p, const int *ne, const double *P0, const double *P1, const int *bdy,
                        double amp, int reference_inverse, double *const *xyz,
                        double *const *xyzp, double *mask, double *const *dxidx, double *jac,
                        double *minnodedist) {
  if (p < 1 || !ne || !P0 || !P1 || !bdy || !xyz || !xyz[0] || !xyz[1] || !dxidx || !jac)
    return GF_ERR_ARG;
  if (bdy[2] == GF_BDY_PERIODIC || bdy[3] == GF_BDY_PERIODIC) return GF_ERR_ARG;
  const int pp[2] = {p, p};
  int rc = gf_box_grid(2, pp, ne, P0, P1, bdy, xyz, xyzp, mask, nullptr, nullptr, nullptr);
  if (rc) return rc;
  const int N = p + 1;
  const double Lx = P1[0] - P0[0], H = P1[1] - P0[1];
  const size_t nel = (size_t)ne[0] * ne[1], ndof = nel * N * N;
  auto move = [&](double *const *c) {
    for (size_t q = 0; q < ndof; q++) {
      const double h = amp * 0.5 * (1.0 - std::cos(2.0 * kPi * (c[0][q] - P0[0]) / Lx));
      c[1][q] = c[1][q] + h * (P1[1] - c[1][q]) / H;
    }
  };
  move(xyz);
  if (xyzp && xyzp[0] && xyzp[1]) move(xyzp);
  rc = gf_deformed_geometry2d(p, nel, xyz, reference_inverse, dxidx, jac);
  if (rc) return rc;
  if (minnodedist) *minnodedist = min_dist2d(xyz, ndof);
  return GF_OK;
}

}  // extern "C"

// Interpolation from the order-p GLL nodes to the order-pq GLL nodes,
// IQP(a, i) = h_i(xq_a) at I[a + (pq+1) i], and the order-pq weights
// (GGrid::init_qdealias, ggrid.ipp:2368-2383).  Host-side.
int gf_dealias_matrix(int p, int pq, double *I, double *wq) {
  if (p < 1 || pq < p || pq > 63 || !I || !wq) return GF_ERR_ARG;
  const int N = p + 1, Q = pq + 1;
  const LDBasis lb = make_ldbasis(p);
  std::vector<double> xq(Q), Dq((size_t)Q * Q);
  if (gf_gll_basis(pq, xq.data(), wq, Dq.data()) != GF_OK) return GF_ERR_ARG;
  for (int a = 0; a < Q; a++)
    for (int i = 0; i < N; i++) I[a + (size_t)Q * i] = eval_basis_m(lb, i, xq[a]);
  return GF_OK;
}
