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
  p.K = m->K32;
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
// Where the fused tails of the CG apply and update kernels keep their group
// results and counters (tail_counters(nblocks) of each), or NULL pointers
// for a separate k_cg_alpha / k_cg_beta launch.
struct CGTail {
  double *gpart = nullptr;
  unsigned *cnt = nullptr;
  double *resid = nullptr, *hmirror = nullptr;  // update tail: as k_cg_beta
};

struct CGApply3Args {
  const void *r, *dinv;
  void *w, *q;
  const void *D, *W, *edx, *ejac, *qf, *mask;
  const void *DR[3], *DC[3];  // pencil copies of D (same precision as T)
  const void *K;              // FP32 stiffness matrix (FP32 applies at N = 8) or NULL
  double p;
  double qs;  // mass-term scalar when qf is NULL (0: no mass term)
  double *scal, *part;
  size_t nelems;
  gf::BoxTopo topo;  // with tmask: mask formed from the box faces (mask NULL)
  int tmask;
  // jdiag: Jacobi diagonal inverse formed in the kernel (box_jacobi_inv)
  int jdiag;
  const void *jA;
  // members of a batched mesh (nseg > 0): element offsets and per-member
  // scalars; partial sums are then indexed by logical block
  const int64_t *seg_eoff;
  int nseg;
  double *sscal;
  // alpha formed at the end of the apply kernel (grid_tail) when tail.cnt
  CGTail tail;
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
static_assert(kUpdBlocks <= GF_RED_MAX_BLOCKS, "update tail counters are sized for GF_RED_MAX_BLOCKS");
constexpr int kBatch = 16;

// ---------------------------------------------------------------------------
// Grid-wide reduction of per-block partials at the end of the kernel that
// produced them, instead of in a second launch.  Blocks are taken in groups
// of kTailGrp consecutive indices.  The last block of a group to finish (a
// device-scope counter per group) reduces the group's partials in index
// order; the last group to finish reduces the group results in index order.
// The total therefore does not depend on the order in which blocks finish.
// Each counter is set back to zero by the block that consumes it, ready for
// the next launch.  part[NV*b + v] holds value v of block b (written by
// thread 0 of block b before the call); mx[v] selects max instead of sum.
// All threads of every block call this; it returns true in thread 0 of the
// one block that holds the totals.  Nothing waits on another block.
// ---------------------------------------------------------------------------
constexpr int kTailGrp = GF_TAIL_GRP;

template <int NV>
__device__ __forceinline__ bool grid_tail(const double *part, double *gpart, unsigned *cnt,
                                          const bool (&mx)[NV], double (&tot)[NV]) {
  __shared__ int s_grp_last, s_all_last;
  const int nb = (int)gridDim.x, ng = (nb + kTailGrp - 1) / kTailGrp;
  const int grp = (int)blockIdx.x / kTailGrp, b0 = grp * kTailGrp;
  const int gs = nb - b0 < kTailGrp ? nb - b0 : kTailGrp;
  // lanes of the first wave (blocks of fewer than 64 threads: all of them;
  // the wave reductions see zeros from absent lanes)
  const int nl = blockDim.x < 64 ? (int)blockDim.x : 64;
  if (threadIdx.x == 0) {
    __threadfence();  // this block's partial before its count
    s_grp_last = atomicAdd(&cnt[grp], 1u) == (unsigned)(gs - 1);
  }
  __syncthreads();
  if (!s_grp_last) return false;
  if (threadIdx.x < 64) {
    __threadfence();  // every partial of the group after the count
    const int lane = (int)threadIdx.x;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = 0.0;
    for (int b = lane; b < gs; b += nl)
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const double x = part[NV * (b0 + b) + q];
        v[q] = mx[q] ? fmax(v[q], x) : v[q] + x;
      }
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = mx[q] ? gf::wave_max(v[q]) : gf::wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NV; q++) gpart[NV * grp + q] = v[q];
      cnt[grp] = 0u;
      __threadfence();
      s_all_last = atomicAdd(&cnt[ng], 1u) == (unsigned)(ng - 1);
    }
  }
  __syncthreads();
  if (!s_all_last) return false;
  if (threadIdx.x < 64) {
    __threadfence();
    const int lane = (int)threadIdx.x;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = 0.0;
    for (int g = lane; g < ng; g += nl)
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const double x = gpart[NV * g + q];
        v[q] = mx[q] ? fmax(v[q], x) : v[q] + x;
      }
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = mx[q] ? gf::wave_max(v[q]) : gf::wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < NV; q++) tot[q] = v[q];
      cnt[ng] = 0u;
    }
  }
  return threadIdx.x == 0;
}

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
  const T *imult;  // 1/multiplicity in the recurrence precision (exact in FP64)
  const void *D, *W, *edx, *ejac;
  const void *G[6];  // T node metrics and lumped mass (meshes with node metrics)
  const void *mass;
  double p;
  double qs;  // mass-term scalar when qf is NULL
  // jdiag: the Jacobi diagonal inverse is formed in the kernels (uniform
  // structured box, scalar p and q) from jA and the node's group size
  int jdiag;
  const T *jA;
};

// Assembled Jacobi diagonal of (p L + q M) at node (i,j,k) of a uniform
// structured box whose group has cnt members: every copy contributes the
// same element diagonal (k_helm_diag's terms: p (sum_d G_dd A(n_d) prod_o W)
// Jac + q W_i W_j W_k Jac), so the sum is cnt times it.
template <typename T>
__device__ __forceinline__ T box_jacobi_inv(const T *__restrict__ A, const T *__restrict__ W,
                                            int i, int j, int k, T g0, T g1, T g2, T jac, T p,
                                            T q, int cnt) {
  T lap = g0 * A[i] * (W[j] * W[k]) * jac;
  lap += g1 * A[j] * (W[i] * W[k]) * jac;
  lap += g2 * A[k] * (W[i] * W[j]) * jac;
  T val = p * lap;
  if (q != T(0)) val += q * (W[k] * (W[j] * W[i]) * jac);
  return T(1) / (val * (T)cnt);
}

// Group size of node (i,j,k) of the element at box coordinates ec.
__device__ __forceinline__ int box_count(const gf::BoxTopo &t, int N, const int (&ec)[3], int i,
                                         int j, int k) {
  const int nc[3] = {i, j, k};
  int c = 1;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool part = (nc[d] == 0 && (ec[d] > 0 || t.per[d])) ||
                      (nc[d] == N - 1 && (ec[d] < t.ne[d] - 1 || t.per[d]));
    c *= part ? 2 : 1;
  }
  return c;
}

// Inverse lumped mass at a node of a box of identical elements whose weight
// table is symmetric (W_i == W_{N-1-i} bit for bit): every member of the
// node's group then has the node's own mass, so the stored inverse mass --
// k_mass_init's ((W_k W_j) W_i) Jac, GGFX Smooth (a sum from 0 over the
// group, then / count) and k_recip's reciprocal -- follows from the group
// size alone, in the same operations and with the same result.
__device__ __forceinline__ double box_imass(double wk, double wj, double wi, double jac, int cnt) {
  const double mn = ((wk * wj) * wi) * jac;
  double s = mn;
#pragma unroll
  for (int q = 1; q < 8; q++)
    if (q < cnt) s = s + mn;
  return 1.0 / (s * gf::box_inv(cnt));
}

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
int launch_update_box_n(gf_mesh *m, CGVecs<T> &v, double *part, const CGTail &tail);

}  // namespace gfs
