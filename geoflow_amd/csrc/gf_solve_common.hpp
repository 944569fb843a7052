// gf_solve_common.hpp -- definitions shared by the solver translation units:
// launch helpers, the device-resident solver scalars, the workgroup-to-element
// mapping of the 3-D element kernels and declarations of their launchers.
// The 3-D kernels themselves are in gf_k3d.hpp and are compiled once per
// element order in the gf_k3d_n*.hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gf_boxdss.hpp"
#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_pen3.hpp"

namespace gfs {

using gf::kThreads;
using gf::Shape;

inline int nblocks(size_t n, int per = kThreads) {
  size_t b = (n + per - 1) / per;
  if (b > 65536) b = 65536;
  return (int)std::max<size_t>(b, 1);
}

// Solver scalars on the device, scal[]: 0 rho, 1 alpha, 2 beta, 3 rnorm,
// 4 rtol, 5 wq, 6 iter, 7 maxit, 8 done, 9 first, 10 volume, 11 n,
// 12 rho of the previous iteration.
enum { S_RHO = 0, S_ALPHA, S_BETA, S_RNORM, S_RTOL, S_WQ, S_ITER, S_MAXIT, S_DONE, S_FIRST,
       S_VOL, S_N, S_RHOM, S_COUNT };

#define GRID_LOOP(g, n)                                                     \
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < (n); \
       g += (size_t)gridDim.x * blockDim.x)

struct StageArgs {
  const double *uin[3];
  const double *kprev[3];
  double bh;
  double *acc[3];
  const double *accsrc[3];
  double ch;
  double *kout[3];
  const double *D, *W, *edx, *ejac, *imass;
  double nu;
  size_t nelems;
};

inline gf::PenD<double> pend_d(gf_mesh *m) {
  gf::PenD<double> p;
  for (int c = 0; c < 3; c++) {
    p.R[c] = m->pDR[c];
    p.C[c] = m->pDC[c];
  }
  return p;
}
inline gf::PenD<float> pend_f(gf_mesh *m) {
  gf::PenD<float> p;
  for (int c = 0; c < 3; c++) {
    p.R[c] = m->pDR32[c];
    p.C[c] = m->pDC32[c];
  }
  return p;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Grids of the element kernels are a multiple of kXCD workgroups and
// are renumbered so that each XCD walks one contiguous range of elements:
// neighbouring elements, whose shared faces the gather kernels read, then
// run on the same XCD close together in time.
constexpr int kXCD = 8;
inline int xcd_grid(size_t nblocks) { return (int)((nblocks + kXCD - 1) / kXCD * kXCD); }
__device__ __forceinline__ size_t xcd_block() {
  const size_t per = gridDim.x / kXCD;
  return (size_t)(blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
}

// Copies the 512 doubles at g (an element, 4 KB) to LDS at l with four
// 16-byte global-to-LDS loads per lane (global_load_lds_dwordx4: lane L's 16
// bytes land at l + 16 L, so l gets g's layout), half the instructions of
// per-column 8-byte loads and no registers holding the data in flight.  l
// must be the same for the whole wave; the compiler waits for the copy before
// the first LDS read after the next barrier.
__device__ __forceinline__ void lds_copy512(const double *__restrict__ g, double *l, int lane) {
#pragma unroll
  for (int q = 0; q < 4; q++)
    __builtin_amdgcn_global_load_lds((const void *)(g + q * 128 + 2 * lane),
                                     (__attribute__((address_space(3))) void *)(l + q * 128), 16,
                                     0, 0);
}

// Grid-stride loop over [0, n) with the blocks renumbered by xcd_block (the
// grid a multiple of kXCD): each XCD walks contiguous ranges.
#define XCD_GRID_LOOP(g, n)                                                 \
  for (size_t g = xcd_block() * blockDim.x + threadIdx.x; g < (n);         \
       g += (size_t)gridDim.x * blockDim.x)

// Thread/element mapping shared by the 3-D element kernels.
template <int N>
struct ColIdx {
  int el, i, j, r;
  size_t e;
  bool active;
  __device__ __forceinline__ ColIdx(size_t nelems) {
    using C = gf::Pen3<N>;
    el = threadIdx.x / C::N2;
    r = threadIdx.x - el * C::N2;
    i = r % N;
    j = r / N;
    size_t eg = xcd_block() * C::EPB + el;
    active = el < C::EPB && eg < nelems;
    e = active ? eg : (nelems - 1);
  }
};

// CG direction update + operator apply (see k_cg_apply for the contract),
// 3-D pencil form; the Dirichlet mask is applied to q here (the mask is
// constant over each coincidence group, so DSS(q) * mask = DSS(q * mask)).

struct CGApply3Args {
  const void *r, *dinv;
  // Jacobi table: elements with deq[e] read dtab (one element's values) in
  // place of dinv; null: dinv everywhere
  const void *dtab;
  const unsigned char *deq;
  void *w, *q;
  const void *D, *W, *edx, *ejac, *qf, *mask;
  const void *DR[3], *DC[3];  // pencil copies of D (same precision as T)
  double p;
  double qs;  // mass-term scalar when qf is NULL (0: no mass term)
  double *scal, *part;
  size_t nelems;
  gf::BoxTopo topo;  // with tmask: mask formed from the box faces (mask NULL)
  int tmask;
  // members of a batched mesh (nseg > 0): element offsets and per-member
  // scalars; partial sums are then indexed by logical block
  const int64_t *seg_eoff;
  int nseg;
  double *sscal;
};

// Solver scalars per member of a batched mesh: kSegScal doubles each.
constexpr int kSegScal = 16;

// Member of element e: the s with eoff[s] <= e < eoff[s+1] (binary search
// over the nseg+1 offsets; e past the end maps to the last member).
__device__ __forceinline__ int seg_of(const int64_t *__restrict__ eoff, int nseg, size_t e) {
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((size_t)eoff[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

inline int col_blocks(gf_mesh *m) {
  int N = m->Niso, N2 = N * N, epb = N2 >= 64 ? 1 : 64 / N2;
  return xcd_grid((m->nelems + epb - 1) / epb);
}

#define GF_DISPATCH_N3(NV, FN, ...)                   \
  [&]() -> int {                                      \
    switch (NV) {                                     \
      case 2: return FN<2>(__VA_ARGS__);              \
      case 3: return FN<3>(__VA_ARGS__);              \
      case 4: return FN<4>(__VA_ARGS__);              \
      case 5: return FN<5>(__VA_ARGS__);              \
      case 6: return FN<6>(__VA_ARGS__);              \
      case 7: return FN<7>(__VA_ARGS__);              \
      case 8: return FN<8>(__VA_ARGS__);              \
      case 9: return FN<9>(__VA_ARGS__);              \
      case 10: return FN<10>(__VA_ARGS__);            \
      default: return GF_ERR_UNSUPPORTED;             \
    }                                                 \
  }()

constexpr int kUpdBlocks = 1024;
constexpr int kBatch = 16;

// alpha = rho / (w.q) from the total w.q.
__device__ __forceinline__ void cg_alpha_finish(double *scal, double wq) {
  scal[S_WQ] = wq;
  scal[S_ALPHA] = scal[S_RHO] / wq;
}

__device__ __forceinline__ double finish_norm(int normtype, double v, const double *scal) {
  switch (normtype) {
    case GF_NORM_INF: return v;
    case GF_NORM_EUC: return sqrt(v / scal[S_N]);
    case GF_NORM_L2: return sqrt(v / scal[S_VOL]);
    case GF_NORM_L1: return v / scal[S_VOL];
    default: return sqrt(v);
  }
}

// rho_new, beta, rnorm, iteration count and the stopping rule of
// gcg.ipp:297-335 (continue while iter < maxit and rnorm > rtol) from the
// totals rho = r.z and nr (norm sum or max).  The scalars are also written to
// hmirror (pinned host memory mapped into the device), from which the host
// polls the stopping flag without a copy.
__device__ __forceinline__ void cg_beta_finish(double *scal, double rho, double nr, int normtype,
                                               double *resid, double *hmirror) {
  const double rhom = scal[S_RHO];
  scal[S_RHOM] = rhom;
  scal[S_RHO] = rho;
  scal[S_BETA] = rho / rhom;
  const double rn = finish_norm(normtype, nr, scal);
  scal[S_RNORM] = rn;
  int it = (int)scal[S_ITER];
  if (resid) resid[it] = rn;
  it++;
  scal[S_ITER] = it;
  scal[S_FIRST] = 0.0;
  if (!(it < (int)scal[S_MAXIT] && rn > scal[S_RTOL])) scal[S_DONE] = 1.0;
  if (hmirror)
    for (int k = 0; k < S_COUNT; k++) hmirror[k] = scal[k];
}


template <typename T>
struct CGVecs {
  T *x, *r, *w, *q;
  const T *dinv, *mask, *qf;
  const T *dtab;               // Jacobi table (CGApply3Args), or null
  const unsigned char *deq;
  const T *imult;  // 1/multiplicity in the recurrence precision (exact in FP64)
  const void *D, *W, *edx, *ejac;
  const void *G[6];  // T node metrics and lumped mass (meshes with node metrics)
  const void *mass;
  double p;
  double qs;  // mass-term scalar when qf is NULL
};




// Launchers of the 3-D element kernels (gf_k3d.hpp), one instantiation per
// element order N = 2..10 in the gf_k3d_n*.hip translation units.
template <int N>
int launch_stage3p(gf_mesh *m, const StageArgs &a);
template <int N>
int launch_stage3b(gf_mesh *m, double *const *uin, double *const *kp, double *const *acc,
                   double *const *kout, double bh, double ch, double nu, int acc_mode,
                   int smooth_v);
template <int N>
int launch_rk_out_box(gf_mesh *m, double *const *u, double *const *acc, double *const *kf,
                      double cl);
template <int N>
int helm3c_d(gf_mesh *m, double p, const double *pf, const double *qf, double qs, const double *u,
             double *uo);
template <int N>
int helm3c_f(gf_mesh *m, float p, const float *qf, float qs, const float *u, float *uo);
template <int N>
int cgapply3c_d(gf_mesh *m, const CGApply3Args &a);
template <int N>
int cgapply3c_f(gf_mesh *m, const CGApply3Args &a);
template <typename T, int NORM, int N>
int launch_update_box_n(gf_mesh *m, CGVecs<T> &v, double *part);

}  // namespace gfs
