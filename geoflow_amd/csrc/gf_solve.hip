// gf_solve.hip -- advection, Burgers RHS and explicit RK steps, the CG solver
// (FP64, and FP32 recurrences with FP64 reliable updates), 2-D element
// kernels and kernel timing.  The 3-D element kernels are in gf_k3d.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "gf_solve_common.hpp"

using namespace gfs;

namespace {

// z = a*x + b*y   (GMTK::saxpby (2), gmtk.ipp:1187-1203); y may alias z.
__global__ void k_axpby(double *__restrict__ z, double a, const double *__restrict__ x, double b,
                        const double *y, size_t n) {
  GRID_LOOP(g, n) z[g] = a * x[g] + b * y[g];
}
// y += a*x
__global__ void k_axpy(double *y, double a, const double *__restrict__ x, size_t n) {
  GRID_LOOP(g, n) y[g] = y[g] + a * x[g];
}
__global__ void k_copy(double *__restrict__ y, const double *__restrict__ x, size_t n) {
  GRID_LOOP(g, n) y[g] = x[g];
}
__global__ void k_zero(double *__restrict__ y, size_t n) {
  GRID_LOOP(g, n) y[g] = 0.0;
}
__global__ void k_nonfinite(const double *__restrict__ x, size_t n, double *flag) {
  bool bad = false;
  GRID_LOOP(g, n) bad |= !isfinite(x[g]);
  if (bad) *flag = 1.0;
}

// GMTK::constrain2sphere (2), in place with the reference's aliasing:
// the updated v0 feeds v1, and both feed v2 (gmtk.ipp:2956-2980).
__global__ void k_constrain2sphere(double *v0, double *v1, double *v2,
                                   const double *__restrict__ X, const double *__restrict__ Y,
                                   const double *__restrict__ Z, size_t n) {
  double x = X[0], y = Y[0], z = Z[0];
  double r2 = x * x + y * y + z * z;
  double ri = 1.0 / r2;
  GRID_LOOP(g, n) {
    x = X[g];
    y = Y[g];
    z = Z[g];
    v0[g] = (v0[g] * (r2 - x * x) - v1[g] * x * y - v2[g] * x * z) * ri;
    v1[g] = (-v0[g] * y * x + v1[g] * (r2 - y * y) - v2[g] * y * z) * ri;
    v2[g] = (-v0[g] * z * x - v1[g] * z * y + v2[g] * (r2 - z * z)) * ri;
  }
}

// ---------------------------------------------------------------------------
// Fused Burgers stage for GE_REGULAR meshes (nonlinear mode, DIM components):
//   u_c   = uin_c + bh * kprev_c                 (stage input, step_b :291-295)
//   acc_c = accsrc_c + ch * kprev_c              (uout accumulation, :306-307)
//   K_c   = Minv * ( -nu L u_c - M sum_j u_j (dXi_j/dX_j) D_j u_c )
// i.e. GBurgers::dudt_impl (gburgers.ipp:247-255) with GAdvect::apply
// (gadvect.ipp:57-99) and GHelmholtz::reg_prod (ghelmholtz.ipp:308-378).
// K is written unassembled; the caller applies the GGFX Smooth.
// ---------------------------------------------------------------------------

template <int DIM, int N>
__global__ __launch_bounds__(kThreads) void k_burgers_stage(StageArgs a) {
  using S = Shape<DIM, N>;
  constexpr int PT = (S::NODES + kThreads - 1) / kThreads;
  __shared__ double su[DIM][S::NODES];
  __shared__ double sg[DIM][S::NODES];
  __shared__ double sD[N * N];
  __shared__ double sW[N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, a.nelems - e0);
  size_t base = e0 * S::NN;
  int cnt = nel * S::NN;
  gf::stage_D(sD, a.D, N);
  if (threadIdx.x < N) sW[threadIdx.x] = a.W[threadIdx.x];
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
      size_t g = base + q;
      double ui = a.uin[c][g];
      double v = ui;
      if (a.kprev[c]) {
        double kp = a.kprev[c][g];
        v = ui + a.bh * kp;
        if (a.acc[c]) a.acc[c][g] = a.accsrc[c][g] + kp * a.ch;
      }
      su[c][q] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    double adv[PT];
    int t = 0;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x, t++) {
      int el = q / S::NN, r = q - el * S::NN, i, j, k;
      gf::node_ijk<DIM, N>(r, i, j, k);
      size_t e = e0 + el;
      double M = gf::node_mass<DIM>(sW, i, j, k, a.ejac[e]);
      double ad = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        double gd = gf::contract<DIM, N, false>(&su[c][el * S::NN], sD, d, i, j, k);
        double f = a.edx[e * DIM + d];
        double phys = gd * f;
        if (d == 0) ad = phys * su[0][q];
        else ad += phys * su[d][q];
        sg[d][q] = M * (gd * (f * f));
      }
      adv[t] = ad * M;
    }
    __syncthreads();
    t = 0;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x, t++) {
      int el = q / S::NN, r = q - el * S::NN, i, j, k;
      gf::node_ijk<DIM, N>(r, i, j, k);
      double L = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; d++) L += gf::contract<DIM, N, true>(&sg[d][el * S::NN], sD, d, i, j, k);
      if (a.nu != 1.0) L *= a.nu;
      double urhs = -1.0 * L + -1.0 * adv[t];
      a.kout[c][base + q] = a.imass[base + q] * urhs;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 3-D regular-element kernels: thread (i,j) of an element owns the node
// column (i,j,0..N-1) and the x/y pencils of gf_pen3.hpp.
// ---------------------------------------------------------------------------
__global__ void k_final_input(double *__restrict__ u3, const double *__restrict__ uin,
                              const double *__restrict__ k, double bh, double *acc,
                              const double *accsrc, double ch, size_t n) {
  GRID_LOOP(g, n) {
    double kv = k[g];
    u3[g] = uin[g] + bh * kv;
    if (acc) acc[g] = accsrc[g] + kv * ch;
  }
}

template <int DIM, int N>
int launch_stage(gf_mesh *m, const StageArgs &a) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  hipLaunchKernelGGL((k_burgers_stage<DIM, N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, a);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// ---------------------------------------------------------------------------
// CG kernels (T = double or float).  Scalars live on the device in
// scal[]: 0 rho, 1 alpha, 2 beta, 3 rnorm, 4 rtol, 5 wq, 6 iter, 7 maxit,
// 8 done, 9 first, 10 volume, 11 n.
// ---------------------------------------------------------------------------

struct CGApplyArgs {
  const void *r;      // T
  const void *dinv;   // T or NULL
  void *w;            // T
  void *q;            // T
  const void *D, *W, *edx, *ejac, *qf;  // T
  const void *G[6];  // T node metrics: 2-D G00, G01, G11; 3-D deformed also G02, G12, G22
  const void *mass;  // T lumped mass per node (meshes with node metrics)
  double p;
  double qs;  // mass-term scalar when qf is NULL (0: no mass term)
  double *scal;
  double *part;
  size_t nelems;
};

// w = beta*w + z (z = Dinv r, or w = z on the first iteration), then
// q = A w on the element (unassembled) and the partial sum of w*q.  Because
// w is continuous and zero where masked, sum_j w_j (DSS q)_j imult_j equals
// sum_j w_j q_j over local nodes, so the dot product of gcg.ipp:307-308 is
// formed here before the DSS.
template <int DIM, int N, typename T>
__global__ __launch_bounds__(kThreads) void k_cg_apply(CGApplyArgs a) {
  using S = Shape<DIM, N>;
  __shared__ T su[S::NODES];
  __shared__ T sg[DIM][S::NODES];
  __shared__ T sD[N * N];
  __shared__ T sW[N];
  __shared__ double sred[kThreads / 64];
  const double *scal = a.scal;
  if (scal[S_DONE] != 0.0) return;
  const T *r = (const T *)a.r, *dinv = (const T *)a.dinv;
  T *w = (T *)a.w, *qv = (T *)a.q;
  const T *D = (const T *)a.D, *W = (const T *)a.W, *edx = (const T *)a.edx,
          *ejac = (const T *)a.ejac, *qf = (const T *)a.qf;
  bool first = scal[S_FIRST] != 0.0;
  T beta = (T)scal[S_BETA];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, a.nelems - e0);
  size_t base = e0 * S::NN;
  int cnt = nel * S::NN;
  gf::stage_D(sD, D, N);
  if (threadIdx.x < N) sW[threadIdx.x] = W[threadIdx.x];
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    size_t g = base + q;
    T z = dinv ? dinv[g] * r[g] : r[g];
    T wn = first ? z : beta * w[g] + (T)1 * z;
    w[g] = wn;
    su[q] = wn;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    size_t e = e0 + el;
    T M = gf::node_mass<DIM>(sW, i, j, k, ejac[e]);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      T gd = gf::contract<DIM, N, false>(su + el * S::NN, sD, d, i, j, k);
      T f = edx[e * DIM + d];
      sg[d][q] = M * (gd * (f * f));
    }
  }
  __syncthreads();
  double dot = 0.0;
  T p = (T)a.p;
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    T acc = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) acc += gf::contract<DIM, N, true>(&sg[d][el * S::NN], sD, d, i, j, k);
    if (p != (T)1) acc *= p;
    if (qf || a.qs != 0.0) {
      T M = gf::node_mass<DIM>(sW, i, j, k, ejac[e0 + el]);
      T t = M * su[q];
      if (qf) t *= qf[base + q];
      else if ((T)a.qs != (T)1) t *= (T)a.qs;
      acc += t;
    }
    qv[base + q] = acc;
    dot += (double)su[q] * (double)acc;
  }
  double bsum = gf::block_reduce<false>(dot, sred);
  if (threadIdx.x == 0) a.part[blockIdx.x] = bsum;
}

// The same for meshes with per-node metrics (GE_2DEMBEDDED and GE_DEFORMED):
// q = p sum_j D_j^T (sum_i G(i,j) D_i w) + q M w in the operation order of
// embed_prod / def_prod (ghelmholtz.ipp:108-191, 205-290, 494-514), as
// k_helm_embed and k_helm_def3 form it; 2-D with G00, G01, G11, 3-D with the
// six vectors G00, G01, G11, G02, G12, G22.  T is the recurrence precision:
// FP64, or FP32 with FP32 copies of G, the mass and D.
template <int DIM, int N, typename T>
__global__ __launch_bounds__(kThreads) void k_cg_apply_gm(CGApplyArgs a) {
  using S = Shape<DIM, N>;
  __shared__ T su[S::NODES];
  __shared__ T sg[DIM][S::NODES];
  __shared__ T sD[N * N];
  __shared__ double sred[kThreads / 64];
  const double *scal = a.scal;
  if (scal[S_DONE] != 0.0) return;
  const T *r = (const T *)a.r, *dinv = (const T *)a.dinv;
  T *w = (T *)a.w, *qv = (T *)a.q;
  const T *qf = (const T *)a.qf, *mass = (const T *)a.mass;
  const T *G0 = (const T *)a.G[0], *G1 = (const T *)a.G[1], *G2 = (const T *)a.G[2];
  const bool first = scal[S_FIRST] != 0.0;
  const T beta = (T)scal[S_BETA];
  const size_t e0 = (size_t)blockIdx.x * S::EPB;
  const int nel = (int)min((size_t)S::EPB, a.nelems - e0);
  const size_t base = e0 * S::NN;
  const int cnt = nel * S::NN;
  gf::stage_D(sD, (const T *)a.D, N);
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const size_t g = base + q;
    const T z = dinv ? dinv[g] * r[g] : r[g];
    const T wn = first ? z : beta * w[g] + (T)1 * z;
    w[g] = wn;
    su[q] = wn;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    const size_t g = base + q;
    const T d0 = gf::contract<DIM, N, false>(su + el * S::NN, sD, 0, i, j, k);
    const T d1 = gf::contract<DIM, N, false>(su + el * S::NN, sD, 1, i, j, k);
    if constexpr (DIM == 2) {
      T g0 = (T)0 + d0 * G0[g];
      g0 += d1 * G1[g];
      T g1 = (T)0 + d0 * G1[g];
      g1 += d1 * G2[g];
      sg[0][q] = g0;
      sg[1][q] = g1;
    } else {
      const T d2 = gf::contract<DIM, N, false>(su + el * S::NN, sD, 2, i, j, k);
      const T a00 = G0[g], a01 = G1[g], a11 = G2[g], a02 = ((const T *)a.G[3])[g],
              a12 = ((const T *)a.G[4])[g], a22 = ((const T *)a.G[5])[g];
      T g0 = (T)0 + d0 * a00;
      g0 += d1 * a01;
      g0 += d2 * a02;
      T g1 = (T)0 + d0 * a01;
      g1 += d1 * a11;
      g1 += d2 * a12;
      T g2 = (T)0 + d0 * a02;
      g2 += d1 * a12;
      g2 += d2 * a22;
      sg[0][q] = g0;
      sg[1][q] = g1;
      sg[DIM - 1][q] = g2;
    }
  }
  __syncthreads();
  double dot = 0.0;
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    T acc = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) acc += gf::contract<DIM, N, true>(&sg[d][el * S::NN], sD, d, i, j, k);
    if (a.p != 1.0) acc *= (T)a.p;
    if (qf || a.qs != 0.0) {
      T t = mass[base + q] * su[q];
      if (qf) t *= qf[base + q];
      else if (a.qs != 1.0) t *= (T)a.qs;
      acc += t;
    }
    qv[base + q] = acc;
    dot += (double)su[q] * (double)acc;
  }
  const double bsum = gf::block_reduce<false>(dot, sred);
  if (threadIdx.x == 0) a.part[blockIdx.x] = bsum;
}

template <typename T>
__global__ void k_dss_cg(T *__restrict__ u, const int *__restrict__ goff,
                         const int *__restrict__ gmem, int ngroups, const double *scal) {
  if (scal[S_DONE] != 0.0) return;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += gridDim.x * blockDim.x) {
    int b = goff[g], e = goff[g + 1];
    if (e - b <= 8) {
      // all member loads issued before the sum (as in k_dss)
      const int cnt = e - b;
      int id[8];
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) id[k] = k < cnt ? gmem[b + k] : 0;
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = k < cnt ? u[id[k]] : T(0);
      T s = T(0);
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < cnt) s = s + v[k];
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < cnt) u[id[k]] = s;
      continue;
    }
    T s = T(0);
    for (int k = b; k < e; k++) s = s + u[gmem[k]];
    for (int k = b; k < e; k++) u[gmem[k]] = s;
  }
}

// alpha = rho / (w.q).  One block of 1024 threads reduces the per-block
// partials of the apply kernel in a fixed order (deterministic).
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void k_cg_alpha(double *scal,
                                                          const double *__restrict__ part, int np) {
  __shared__ double sred[kRedThreads / 64];
  if (scal[S_DONE] != 0.0) return;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * kRedThreads < np; i += 4 * kRedThreads) {
    v0 += part[i];
    v1 += part[i + kRedThreads];
    v2 += part[i + 2 * kRedThreads];
    v3 += part[i + 3 * kRedThreads];
  }
  for (; i < np; i += kRedThreads) v0 += part[i];
  double v = (v0 + v1) + (v2 + v3);
  v = gf::wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sred[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double wq = sred[0];
    for (int w = 1; w < kRedThreads / 64; w++) wq += sred[w];
    cg_alpha_finish(scal, wq);
  }
}

// x += alpha w ; r -= alpha mask q ; partials of r.z.imult and of the norm.
template <typename T, int NORM>
__global__ void k_cg_update(T *__restrict__ x, const T *__restrict__ w, T *__restrict__ r,
                            const T *__restrict__ qv, const T *__restrict__ mask,
                            const T *__restrict__ dinv, const T *__restrict__ imult,
                            const double *__restrict__ mass, double *scal,
                            double *__restrict__ part, size_t n) {
  __shared__ double sred[kThreads / 64];
  if (scal[S_DONE] != 0.0) return;
  T alpha = (T)scal[S_ALPHA];
  double rho = 0.0, nrm = 0.0;
  GRID_LOOP(g, n) {
    const T dv = dinv ? dinv[g] : T(1);
    const double im = (double)imult[g];
    T xv = (T)1 * x[g] + alpha * w[g];
    T qm = mask ? qv[g] * mask[g] : qv[g];
    T rv = (T)1 * r[g] + (-alpha) * qm;
    x[g] = xv;
    r[g] = rv;
    T z = dinv ? dv * rv : rv;
    rho += (double)rv * (double)z * (double)im;
    double rd = (double)rv;
    if (NORM == GF_NORM_INF) nrm = fmax(nrm, fabs(rd));
    else if (NORM == GF_NORM_EUC) nrm += rd * rd;
    else if (NORM == GF_NORM_L2) nrm += mass[g] * rd * rd;
    else if (NORM == GF_NORM_L1) nrm += mass[g] * fabs(rd);
    else nrm += rd * rd * (double)im;
  }
  double a = gf::block_reduce<false>(rho, sred);
  double b = NORM == GF_NORM_INF ? gf::block_reduce<true>(nrm, sred)
                                 : gf::block_reduce<false>(nrm, sred);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// rho_new, beta, rnorm, iteration count and the stopping rule
// (cg_beta_finish) from the update kernel's (rho, norm) partials.
__global__ void k_cg_beta(double *scal, const double *__restrict__ part, int np, int normtype,
                          double *resid, double *hmirror) {
  __shared__ double sred[kRedThreads / 64];
  if (scal[S_DONE] != 0.0) return;
  double v = 0.0, w = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    v += part[2 * i];
    w = normtype == GF_NORM_INF ? fmax(w, part[2 * i + 1]) : w + part[2 * i + 1];
  }
  double rho = gf::block_reduce<false>(v, sred);
  double nr = normtype == GF_NORM_INF ? gf::block_reduce<true>(w, sred)
                                      : gf::block_reduce<false>(w, sred);
  if (threadIdx.x == 0) cg_beta_finish(scal, rho, nr, normtype, resid, hmirror);
}

// ---- CG on a batched mesh: every member its own scalars ------------------
// Update blocks are assigned to members in proportion to their size
// (boff[s] .. boff[s+1]); each block works its member's node range only.
__device__ __forceinline__ int seg_of_block(const int *__restrict__ boff, int nseg, int b) {
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (boff[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename T, int NORM>
__global__ void k_cg_update_seg(T *__restrict__ x, const T *__restrict__ w, T *__restrict__ r,
                                const T *__restrict__ qv, const T *__restrict__ mask,
                                const T *__restrict__ dinv, const T *__restrict__ imult,
                                const double *__restrict__ mass, const double *sscal,
                                const int64_t *__restrict__ eoff, const int *__restrict__ boff,
                                int nseg, size_t nn, double *__restrict__ part) {
  __shared__ double sred[kThreads / 64];
  const int s = seg_of_block(boff, nseg, blockIdx.x);
  const double *scal = sscal + kSegScal * s;
  double rho = 0.0, nrm = 0.0;
  if (scal[S_DONE] == 0.0) {
    const T alpha = (T)scal[S_ALPHA];
    const size_t n0 = (size_t)eoff[s] * nn, n1 = (size_t)eoff[s + 1] * nn;
    const size_t lb = blockIdx.x - boff[s], nbs = boff[s + 1] - boff[s];
    for (size_t g = n0 + lb * blockDim.x + threadIdx.x; g < n1; g += nbs * blockDim.x) {
      const T dv = dinv ? dinv[g] : T(1);
      const double im = (double)imult[g];
      T xv = (T)1 * x[g] + alpha * w[g];
      T qm = mask ? qv[g] * mask[g] : qv[g];
      T rv = (T)1 * r[g] + (-alpha) * qm;
      x[g] = xv;
      r[g] = rv;
      T z = dinv ? dv * rv : rv;
      rho += (double)rv * (double)z * im;
      double rd = (double)rv;
      if (NORM == GF_NORM_INF) nrm = fmax(nrm, fabs(rd));
      else if (NORM == GF_NORM_EUC) nrm += rd * rd;
      else if (NORM == GF_NORM_L2) nrm += mass[g] * rd * rd;
      else if (NORM == GF_NORM_L1) nrm += mass[g] * fabs(rd);
      else nrm += rd * rd * im;
    }
  }
  double a = gf::block_reduce<false>(rho, sred);
  double b = NORM == GF_NORM_INF ? gf::block_reduce<true>(nrm, sred)
                                 : gf::block_reduce<false>(nrm, sred);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// alpha of every member: one wave per member, over the member's apply
// partials (logical blocks eoff[s]/epb .. eoff[s+1]/epb).
__global__ __launch_bounds__(1024) void k_cg_alpha_seg(double *sscal, const double *__restrict__ part,
                                                       const int64_t *__restrict__ eoff, int nseg,
                                                       int epb) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int s = wid; s < nseg; s += nw) {
    double *sc = sscal + kSegScal * s;
    if (sc[S_DONE] != 0.0) continue;
    const int64_t b0 = eoff[s] / epb, b1 = eoff[s + 1] / epb;
    double v = 0.0;
    for (int64_t i = b0 + lane; i < b1; i += 64) v += part[i];
    v = gf::wave_sum(v);
    if (lane == 0) {
      sc[S_WQ] = v;
      sc[S_ALPHA] = sc[S_RHO] / v;
    }
  }
}

// rho, beta, norm, iteration count and stopping test of every member (the
// rule of k_cg_beta per member), one wave per member; then the batch is
// done when every member is, which scal[S_DONE] and the host mirror record.
__global__ __launch_bounds__(1024) void k_cg_beta_seg(double *scal, double *sscal,
                                                      const double *__restrict__ part,
                                                      const int *__restrict__ boff, int nseg,
                                                      int normtype, double *hmirror) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int s = wid; s < nseg; s += nw) {
    double *sc = sscal + kSegScal * s;
    if (sc[S_DONE] != 0.0) continue;
    double v = 0.0, w = 0.0;
    for (int i = boff[s] + lane; i < boff[s + 1]; i += 64) {
      v += part[2 * i];
      w = normtype == GF_NORM_INF ? fmax(w, part[2 * i + 1]) : w + part[2 * i + 1];
    }
    v = gf::wave_sum(v);
    w = normtype == GF_NORM_INF ? gf::wave_max(w) : gf::wave_sum(w);
    if (lane == 0) {
      const double rhom = sc[S_RHO];
      sc[S_RHOM] = rhom;
      sc[S_RHO] = v;
      sc[S_BETA] = v / rhom;
      const double rn = finish_norm(normtype, w, sc);
      sc[S_RNORM] = rn;
      const int it = (int)sc[S_ITER] + 1;
      sc[S_ITER] = it;
      sc[S_FIRST] = 0.0;
      if (!(it < (int)sc[S_MAXIT] && rn > sc[S_RTOL])) sc[S_DONE] = 1.0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    bool all = true;
    for (int s = 0; s < nseg; s++) all = all && sscal[kSegScal * s + S_DONE] != 0.0;
    scal[S_DONE] = all ? 1.0 : 0.0;
    if (hmirror) hmirror[S_DONE] = scal[S_DONE];
  }
}

// Set-up reductions per member (rho = sum r z imult and the norm's sum),
// blocks assigned as in k_cg_update_seg; k_setup_fin_seg finishes them,
// out[2s], out[2s+1].
template <typename T>
__global__ void k_setup_red_seg(const T *__restrict__ r, const T *__restrict__ z,
                                const T *__restrict__ imult, const double *__restrict__ mass,
                                int normtype, const int64_t *__restrict__ eoff,
                                const int *__restrict__ boff, int nseg, size_t nn,
                                double *__restrict__ part) {
  __shared__ double sred[kThreads / 64];
  const int s = seg_of_block(boff, nseg, blockIdx.x);
  const size_t n0 = (size_t)eoff[s] * nn, n1 = (size_t)eoff[s + 1] * nn;
  const size_t lb = blockIdx.x - boff[s], nbs = boff[s + 1] - boff[s];
  double rho = 0.0, nrm = 0.0;
  for (size_t g = n0 + lb * blockDim.x + threadIdx.x; g < n1; g += nbs * blockDim.x) {
    double rv = (double)r[g];
    if (z) rho += rv * (double)z[g] * (double)imult[g];
    if (normtype == GF_NORM_INF) nrm = fmax(nrm, fabs(rv));
    else if (normtype == GF_NORM_EUC) nrm += rv * rv;
    else if (normtype == GF_NORM_L2) nrm += mass[g] * rv * rv;
    else if (normtype == GF_NORM_L1) nrm += mass[g] * fabs(rv);
    else nrm += rv * rv * (double)imult[g];
  }
  double a = gf::block_reduce<false>(rho, sred);
  double b = normtype == GF_NORM_INF ? gf::block_reduce<true>(nrm, sred)
                                     : gf::block_reduce<false>(nrm, sred);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}
__global__ __launch_bounds__(1024) void k_setup_fin_seg(const double *__restrict__ part,
                                                        const int *__restrict__ boff, int nseg,
                                                        int normtype, double *out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int s = wid; s < nseg; s += nw) {
    double v = 0.0, w = 0.0;
    for (int i = boff[s] + lane; i < boff[s + 1]; i += 64) {
      v += part[2 * i];
      w = normtype == GF_NORM_INF ? fmax(w, part[2 * i + 1]) : w + part[2 * i + 1];
    }
    v = gf::wave_sum(v);
    w = normtype == GF_NORM_INF ? gf::wave_max(w) : gf::wave_sum(w);
    if (lane == 0) {
      out[2 * s] = v;
      out[2 * s + 1] = w;
    }
  }
}

// Generic reductions used at solver set-up: partial (sum_i a*b*c) and norm.
template <typename T>
__global__ void k_setup_red(const T *__restrict__ r, const T *__restrict__ z,
                            const T *__restrict__ imult, const double *__restrict__ mass,
                            int normtype, double *__restrict__ part, size_t n) {
  __shared__ double sred[kThreads / 64];
  double rho = 0.0, nrm = 0.0;
  GRID_LOOP(g, n) {
    double rv = (double)r[g];
    if (z) rho += rv * (double)z[g] * (double)imult[g];
    if (normtype == GF_NORM_INF) nrm = fmax(nrm, fabs(rv));
    else if (normtype == GF_NORM_EUC) nrm += rv * rv;
    else if (normtype == GF_NORM_L2) nrm += mass[g] * rv * rv;
    else if (normtype == GF_NORM_L1) nrm += mass[g] * fabs(rv);
    else nrm += rv * rv * (double)imult[g];
  }
  double a = gf::block_reduce<false>(rho, sred);
  double b = normtype == GF_NORM_INF ? gf::block_reduce<true>(nrm, sred)
                                     : gf::block_reduce<false>(nrm, sred);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}
__global__ void k_setup_fin(const double *__restrict__ part, int np, int normtype, double *out) {
  __shared__ double sred[kThreads / 64];
  double v = 0.0, w = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    v += part[2 * i];
    w = normtype == GF_NORM_INF ? fmax(w, part[2 * i + 1]) : w + part[2 * i + 1];
  }
  double a = gf::block_reduce<false>(v, sred);
  double b = normtype == GF_NORM_INF ? gf::block_reduce<true>(w, sred)
                                     : gf::block_reduce<false>(w, sred);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

template <typename T>
__global__ void k_mul_t(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ out,
                        size_t n) {
  GRID_LOOP(g, n) out[g] = a[g] * b[g];
}
__global__ void k_fill_f32(float *__restrict__ o, float v, size_t n) {
  GRID_LOOP(g, n) o[g] = v;
}
__global__ void k_to_f32(float *__restrict__ o, const double *__restrict__ i, double s, size_t n) {
  GRID_LOOP(g, n) o[g] = (float)(i[g] * s);
}
__global__ void k_from_f32(double *__restrict__ o, const float *__restrict__ i, size_t n) {
  GRID_LOOP(g, n) o[g] = (double)i[g];
}
__global__ void k_add_f32(double *__restrict__ x, const float *__restrict__ d, double s, size_t n) {
  GRID_LOOP(g, n) x[g] = x[g] + s * (double)d[g];
}
__global__ void k_residual(double *__restrict__ r, const double *__restrict__ b,
                           const double *__restrict__ ax, size_t n) {
  GRID_LOOP(g, n) r[g] = b[g] - ax[g];
}

// Element diagonal of (p L + q M) for regular elements, then assembled by the
// caller with a DSS Sum:
//   p * sum_d G_dd Jac prod_{o!=d} W(i_o) sum_m D(m,i_d)^2 W(m)  +  q M
template <int DIM, int N>
__global__ void k_helm_diag(double *__restrict__ d, const double *__restrict__ D,
                            const double *__restrict__ W, const double *__restrict__ edx,
                            const double *__restrict__ ejac, double p,
                            const double *__restrict__ qf, double qs, size_t ndof) {
  constexpr int NN = Shape<DIM, N>::NN;
  GRID_LOOP(g, ndof) {
    size_t e = g / NN;
    int r = (int)(g - e * NN), idx[3];
    gf::node_ijk<DIM, N>(r, idx[0], idx[1], idx[2]);
    double jac = ejac[e];
    double lap = 0.0;
    for (int dd = 0; dd < DIM; dd++) {
      double s = 0.0;
      for (int m = 0; m < N; m++) {
        double Dm = D[m + idx[dd] * N];
        s += Dm * Dm * W[m];
      }
      double other = 1.0;
      for (int o = 0; o < DIM; o++)
        if (o != dd) other *= W[idx[o]];
      double f = edx[e * DIM + dd];
      double G = f * f;
      lap += G * s * other * jac;
    }
    double val = p * lap;
    if (qf) val += qf[g] * gf::node_mass<DIM>(W, idx[0], idx[1], idx[2], jac);
    else if (qs != 0.0) val += qs * gf::node_mass<DIM>(W, idx[0], idx[1], idx[2], jac);
    d[g] = val;
  }
}
// Element diagonal of (p L + q M) on meshes with node metrics (embedded and
// deformed): the operator sum_{a,b} D_a^T diag(G_ab) D_b has diagonal
//   sum_a sum_l D(l, n_a)^2 G_aa(line_a(n, l)) + 2 sum_{a<b} D(n_a,n_a) D(n_b,n_b) G_ab(n)
// with G_ab in slot b(b+1)/2 + a (G00, G01, G11, G02, G12, G22).
struct DiagG {
  const double *G[6];
};
template <int DIM, int N>
__global__ void k_helm_diag_g(double *__restrict__ d, const double *__restrict__ D, DiagG gm,
                              const double *__restrict__ mass, double p,
                              const double *__restrict__ qf, double qs, size_t ndof) {
  constexpr int NN = Shape<DIM, N>::NN;
  GRID_LOOP(g, ndof) {
    size_t e = g / NN;
    int r = (int)(g - e * NN), idx[3];
    gf::node_ijk<DIM, N>(r, idx[0], idx[1], idx[2]);
    const size_t stride[3] = {1, (size_t)N, (size_t)N * N};
    double lap = 0.0;
    for (int a = 0; a < DIM; a++) {
      const double *Ga = gm.G[a * (a + 1) / 2 + a];
      const size_t line = g - (size_t)idx[a] * stride[a];
      double s = 0.0;
      for (int l = 0; l < N; l++) {
        double Dl = D[l + idx[a] * N];
        s += Dl * Dl * Ga[line + (size_t)l * stride[a]];
      }
      lap += s;
    }
    for (int a = 0; a < DIM; a++)
      for (int b = a + 1; b < DIM; b++)
        lap += 2.0 * D[idx[a] * (N + 1)] * D[idx[b] * (N + 1)] * gm.G[b * (b + 1) / 2 + a][g];
    double val = p * lap;
    if (qf) val += qf[g] * mass[g];
    else if (qs != 0.0) val += qs * mass[g];
    d[g] = val;
  }
}
// dinv = 1/d; *bad = 1 when some entry of d is not positive (or NaN)
__global__ void k_recip_pos(double *__restrict__ a, double *__restrict__ bad, size_t n) {
  bool any = false;
  GRID_LOOP(g, n) {
    const double d = a[g];
    any = any || !(d > 0.0);
    a[g] = 1.0 / d;
  }
  if (any) *bad = 1.0;
}

template <int DIM, int N>
int launch_diag(gf_mesh *m, double p, const double *qf, double qs, double *d) {
  if (m->gtype != GF_GE_REGULAR) {
    DiagG gm;
    for (int c = 0; c < 6; c++) gm.G[c] = m->G[c];
    hipLaunchKernelGGL((k_helm_diag_g<DIM, N>), dim3(nblocks(m->ndof)), dim3(kThreads), 0,
                       m->ctx->stream, d, m->D[0], gm, m->mass, p, qf, qs, m->ndof);
    GF_LAUNCH_CHECK(m->ctx);
    return GF_OK;
  }
  hipLaunchKernelGGL((k_helm_diag<DIM, N>), dim3(nblocks(m->ndof)), dim3(kThreads), 0,
                     m->ctx->stream, d, m->D[0], m->W[0], m->edx, m->ejac, p, qf, qs, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int DIM, int N, typename T>
int launch_cg_apply(gf_mesh *m, const CGApplyArgs &a) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  if (m->gtype != GF_GE_REGULAR) {
    hipLaunchKernelGGL((k_cg_apply_gm<DIM, N, T>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, a);
    GF_LAUNCH_CHECK(m->ctx);
    return GF_OK;
  }
  hipLaunchKernelGGL((k_cg_apply<DIM, N, T>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, a);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}
template <int DIM, int N>
int launch_cg_apply_d(gf_mesh *m, const CGApplyArgs &a) { return launch_cg_apply<DIM, N, double>(m, a); }
template <int DIM, int N>
int launch_cg_apply_f(gf_mesh *m, const CGApplyArgs &a) { return launch_cg_apply<DIM, N, float>(m, a); }

// FP32 element Helmholtz (used for the FP32 operator entry point).
template <int DIM, int N>
__global__ __launch_bounds__(kThreads) void k_helm_f32(const float *__restrict__ u,
                                                       float *__restrict__ uo,
                                                       const float *__restrict__ D,
                                                       const float *__restrict__ W,
                                                       const float *__restrict__ edx,
                                                       const float *__restrict__ ejac, float p,
                                                       const float *__restrict__ qf, float qs,
                                                       size_t nelems) {
  using S = Shape<DIM, N>;
  __shared__ float su[S::NODES];
  __shared__ float sg[DIM][S::NODES];
  __shared__ float sD[N * N];
  __shared__ float sW[N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  int cnt = nel * S::NN;
  gf::stage_D(sD, D, N);
  if (threadIdx.x < N) sW[threadIdx.x] = W[threadIdx.x];
  gf::stage(su, u, base, cnt);
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    size_t e = e0 + el;
    float M = gf::node_mass<DIM>(sW, i, j, k, ejac[e]);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      float gd = gf::contract<DIM, N, false>(su + el * S::NN, sD, d, i, j, k);
      float f = edx[e * DIM + d];
      sg[d][q] = M * (gd * (f * f));
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int el = q / S::NN, rr = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(rr, i, j, k);
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; d++) acc += gf::contract<DIM, N, true>(&sg[d][el * S::NN], sD, d, i, j, k);
    if (p != 1.f) acc *= p;
    if (qf || qs != 0.f) {
      float M = gf::node_mass<DIM>(sW, i, j, k, ejac[e0 + el]);
      float t = M * su[q];
      if (qf) t *= qf[base + q];
      else if (qs != 1.f) t *= qs;
      acc += t;
    }
    uo[base + q] = acc;
  }
}
template <int DIM, int N>
int launch_helm_f32(gf_mesh *m, float p, const float *qf, float qs, const float *u, float *uo) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  hipLaunchKernelGGL((k_helm_f32<DIM, N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, uo,
                     m->D32, m->W32, m->edx32, m->ejac32, p, qf, qs, m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

}  // namespace

int gf_helm3c(gf_mesh *m, double p, const double *pf, const double *qf, double qs,
              const double *u, double *uo) {
  return GF_DISPATCH_N3(m->Niso, helm3c_d, m, p, pf, qf, qs, u, uo);
}

int gf_axpy_t(gf_mesh *m, double a, const double *x, double *y) {
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, y, a, x,
                     m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

static int axpby(gf_mesh *m, double *z, double a, const double *x, double b, const double *y) {
  hipLaunchKernelGGL(k_axpby, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, z, a, x,
                     b, y, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}
static int copy(gf_mesh *m, double *y, const double *x) {
  GF_HIP(m->ctx, hipMemcpyAsync(y, x, m->ndof * sizeof(double), hipMemcpyDeviceToDevice,
                                m->ctx->stream));
  return GF_OK;
}

#define RC(x)            \
  do {                   \
    int _rc = (x);       \
    if (_rc) return _rc; \
  } while (0)

extern "C" {

int gf_helmholtz_apply_f32(gf_mesh *m, float p_scalar, float q_scalar, const float *q_field,
                           const float *u, float *uo) {
  if (!m || !u || !uo) return GF_ERR_ARG;
  if (m->gtype != GF_GE_REGULAR || !m->Niso)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "FP32 Helmholtz: regular isotropic meshes");
  if (m->dim == 3 && m->dconst)
    return GF_DISPATCH_N3(m->Niso, helm3c_f, m, p_scalar, q_field, q_scalar, u, uo);
  return GF_DISPATCH_N(m->dim, m->Niso, launch_helm_f32, m, p_scalar, q_field, q_scalar, u, uo);
}

// GAdvect::apply (gadvect.ipp:57-99): po = deriv(p,1)*u0 ; po += deriv(p,j+1)*uj ;
// po *= mass.
int gf_advect(gf_mesh *m, const double *p, int nvel, const double *const *u, double *po) {
  if (!m || !p || !u || !po || nvel < 1) return GF_ERR_ARG;
  double *t = (double *)gf_ws(m, WS_TMP2, m->ndof * sizeof(double));
  if (!t) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  if (u[0]) {
    RC(gf_deriv(m, 1, p, po));
    RC(gf_mul_t(m, po, u[0], po));
  } else {
    GF_HIP(m->ctx, hipMemsetAsync(po, 0, m->ndof * sizeof(double), m->ctx->stream));
  }
  for (int j = 1; j < nvel; j++) {
    if (!u[j]) continue;
    RC(gf_deriv(m, j + 1, p, t));
    RC(gf_mul_t(m, t, u[j], t));
    RC(gf_axpy_t(m, 1.0, t, po));
  }
  return gf_mul_t(m, po, m->mass, po);
}

// GBurgers::dudt_impl (gburgers.ipp:202-257), composed from the operators.
int gf_burgers_rhs(gf_mesh *m, int mode, double nu, int nstate, const double *const *u,
                   double *const *dudt) {
  if (!m || !u || !dudt || nstate < 1 || mode < 0 || mode > 2) return GF_ERR_ARG;
  for (int k = 0; k < nstate; k++)
    if (!u[k]) return gf_set_err(m->ctx, GF_ERR_ARG, "NULL state component");
  for (int k = 0; k < (mode == 1 ? 1 : nstate); k++)
    if (!dudt[k]) return gf_set_err(m->ctx, GF_ERR_ARG, "NULL output component");
  double *urhs = (double *)gf_ws(m, WS_TMP3, m->ndof * sizeof(double));
  if (!urhs) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  if (mode == 2) {
    for (int k = 0; k < nstate; k++) {
      RC(gf_helm_t(m, nu, nullptr, nullptr, 0.0, u[k], urhs));
      RC(axpby(m, urhs, -1.0, urhs, 0.0, urhs));
      RC(gf_mul_t(m, m->imass, urhs, dudt[k]));
    }
    return GF_OK;
  }
  if (mode == 1) {
    RC(gf_advect(m, u[0], nstate - 1, u + 1, dudt[0]));
    RC(gf_helm_t(m, nu, nullptr, nullptr, 0.0, u[0], urhs));
    RC(axpby(m, urhs, -1.0, urhs, -1.0, dudt[0]));
    return gf_mul_t(m, m->imass, urhs, dudt[0]);
  }
  if (nstate < m->dim) return gf_set_err(m->ctx, GF_ERR_ARG, "nonlinear Burgers needs dim components");
  for (int k = 0; k < m->dim; k++) {
    RC(gf_advect(m, u[k], nstate, u, dudt[k]));
    RC(gf_helm_t(m, nu, nullptr, nullptr, 0.0, u[k], urhs));
    RC(axpby(m, urhs, -1.0, urhs, -1.0, dudt[k]));
    RC(gf_mul_t(m, m->imass, urhs, dudt[k]));
  }
  // dudt_impl forms the RHS of the first GDIM components only
  // (gburgers.ipp:247); the third velocity component of a sphere grid gets
  // a zero RHS and changes only through constrain2sphere.
  for (int k = m->dim; k < nstate; k++) {
    hipLaunchKernelGGL(k_zero, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, dudt[k],
                       m->ndof);
    GF_LAUNCH_CHECK(m->ctx);
  }
  return GF_OK;
}

}  // extern "C"

namespace {

void butcher(int order, double *alpha, double *beta /*4x4*/, double *c) {
  for (int i = 0; i < 4; i++) {
    alpha[i] = 0;
    c[i] = 0;
    for (int j = 0; j < 4; j++) beta[i * 4 + j] = 0;
  }
  // GButcherRK::computeCoeffs (gbutcherrk.ipp:80-127)
  if (order == 1) {
    c[0] = 1.0;
  } else if (order == 2) {
    c[0] = 0.5; alpha[1] = 1.0; c[1] = 0.5; beta[1 * 4 + 0] = 1.0;
  } else if (order == 3) {
    c[0] = 0.25; alpha[1] = 1.0 / 3.0; alpha[2] = 2.0 / 3.0; c[2] = 0.75;
    beta[1 * 4 + 0] = 1.0 / 3.0; beta[2 * 4 + 1] = 2.0 / 3.0;
  } else {
    c[0] = 1.0 / 6.0; alpha[1] = 0.5; c[1] = 1.0 / 3.0; alpha[2] = 0.5; c[2] = 1.0 / 3.0;
    alpha[3] = 1.0; c[3] = 1.0 / 6.0;
    beta[1 * 4 + 0] = 0.5; beta[2 * 4 + 1] = 0.5; beta[3 * 4 + 2] = 1.0;
  }
}

int constrain(gf_mesh *m, double *const *v, int nv) {
  if (m->gtype != GF_GE_2DEMBEDDED || nv < 3) return GF_OK;
  if (!m->xyz[0] || !m->xyz[1] || !m->xyz[2])
    return gf_set_err(m->ctx, GF_ERR_ARG, "constrain2sphere needs node coordinates");
  hipLaunchKernelGGL(k_constrain2sphere, dim3(nblocks(m->ndof)), dim3(kThreads), 0,
                     m->ctx->stream, v[0], v[1], v[2], m->xyz[0], m->xyz[1], m->xyz[2], m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// Fused RK step for regular isotropic meshes, nonlinear mode.
int rk_step_fused(gf_mesh *m, double nu, int order, double dt, double *const *u) {
  int nc = m->dim;
  size_t n = m->ndof, bytes = n * sizeof(double);
  double alpha[4], beta[16], c[4];
  butcher(order, alpha, beta, c);
  double *acc[3], *K[4][3], *u3[3];
  for (int k = 0; k < nc; k++) {
    acc[k] = (double *)gf_ws(m, WS_RK_BEGIN + k, bytes);
    u3[k] = (double *)gf_ws(m, WS_RK_BEGIN + 3 + k, bytes);
    for (int s = 0; s < 2; s++) K[s][k] = (double *)gf_ws(m, WS_RK_BEGIN + 6 + 3 * s + k, bytes);
    if (!acc[k] || !u3[k] || !K[0][k] || !K[1][k]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  }
  double h = dt;
  int ns = order;
  StageArgs a;
  memset(&a, 0, sizeof(a));
  a.D = m->D[0];
  a.W = m->W[0];
  a.edx = m->edx;
  a.ejac = m->ejac;
  a.imass = m->imass;
  a.nu = nu;
  a.nelems = m->nelems;
  // K_m lives in ping-pong buffers: K[m%2]
  for (int s = 0; s < ns - 1; s++) {
    for (int k = 0; k < nc; k++) {
      a.uin[k] = u[k];
      a.kprev[k] = s > 0 ? K[(s - 1) % 2][k] : nullptr;
      a.acc[k] = s > 0 ? acc[k] : nullptr;
      a.accsrc[k] = s == 1 ? u[k] : acc[k];
      a.kout[k] = K[s % 2][k];
    }
    a.bh = s > 0 ? beta[s * 4 + (s - 1)] * h : 0.0;
    a.ch = s > 0 ? c[s - 1] * h : 0.0;
    RC(m->dim == 3 ? GF_DISPATCH_N3(m->Niso, launch_stage3p, m, a)
                   : GF_DISPATCH_N(m->dim, m->Niso, launch_stage, m, a));
    for (int k = 0; k < nc; k++) RC(gf_dss_t(m, GF_DSS_SMOOTH, K[s % 2][k]));
  }
  // final stage input
  for (int k = 0; k < nc; k++) {
    if (ns >= 2) {
      const double *kp = K[(ns - 2) % 2][k];
      double *accd = acc[k];
      const double *accs = ns == 2 ? u[k] : acc[k];
      hipLaunchKernelGGL(k_final_input, dim3(nblocks(n)), dim3(kThreads), 0, m->ctx->stream, u3[k],
                         u[k], kp, beta[(ns - 1) * 4 + (ns - 2)] * h, accd, accs, c[ns - 2] * h, n);
      GF_LAUNCH_CHECK(m->ctx);
    } else {
      RC(copy(m, u3[k], u[k]));
      RC(copy(m, acc[k], u[k]));
    }
    RC(gf_dss_t(m, GF_DSS_SMOOTH, u3[k]));
  }
  for (int k = 0; k < nc; k++) {
    a.uin[k] = u3[k];
    a.kprev[k] = nullptr;
    a.acc[k] = nullptr;
    a.accsrc[k] = nullptr;
    a.kout[k] = K[(ns - 1) % 2][k];
  }
  a.bh = 0.0;
  a.ch = 0.0;
  RC(m->dim == 3 ? GF_DISPATCH_N3(m->Niso, launch_stage3p, m, a)
                 : GF_DISPATCH_N(m->dim, m->Niso, launch_stage, m, a));
  double cl = c[ns - 1] * h;
  for (int k = 0; k < nc; k++) {
    double *kf = K[(ns - 1) % 2][k];
    RC(gf_dss_t(m, GF_DSS_SMOOTH, kf));
    // uout = acc + kf*(c h)
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(kThreads), 0, m->ctx->stream, u[k], 1.0,
                       acc[k], cl, kf, n);
    GF_LAUNCH_CHECK(m->ctx);
    RC(gf_dss_t(m, GF_DSS_SMOOTH, u[k]));
  }
  return GF_OK;
}

// RK steps on a structured box (order >= 2): the same arithmetic as
// rk_step_fused, with every Smooth done by gather inside the stage kernels
// and the output kernel.  Per step: `order` stage kernels and one output
// kernel, no DSS passes.  Between consecutive steps the output kernel and the
// next step's first stage are one kernel (acc_mode 3), which computes the
// output value with the output kernel's operations, stores it and continues
// with it as the stage input, so n steps give the same bits as n single
// steps.
int rk_steps_box(gf_mesh *m, double nu, int order, double dt, int nsteps, double *const *u) {
  size_t bytes = m->ndof * sizeof(double);
  double alpha[4], beta[16], c[4];
  butcher(order, alpha, beta, c);
  double *acc[3], *K[2][3];
  for (int k = 0; k < 3; k++) {
    acc[k] = (double *)gf_ws(m, WS_RK_BEGIN + k, bytes);
    for (int s = 0; s < 2; s++) K[s][k] = (double *)gf_ws(m, WS_RK_BEGIN + 6 + 3 * s + k, bytes);
    if (!acc[k] || !K[0][k] || !K[1][k]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  }
  const double h = dt;
  const int ns = order;
  int kb = 0;  // buffer of the current step's first stage derivative
  bool have_k0 = false;
  for (int n = 0; n < nsteps; n++) {
    // stage 1 from the state itself (unless the previous step's output
    // kernel formed it)
    if (!have_k0)
      RC(GF_DISPATCH_N3(m->Niso, launch_stage3b, m, u, (double *const *)nullptr,
                        (double *const *)nullptr, K[kb], 0.0, 0.0, nu, 0, 0));
    // middle stages and the final stage (whose input is smoothed)
    for (int s = 1; s < ns; s++) {
      const bool last = s == ns - 1;
      RC(GF_DISPATCH_N3(m->Niso, launch_stage3b, m, u, K[(kb + s - 1) % 2], acc, K[(kb + s) % 2],
                        beta[s * 4 + (s - 1)] * h, c[s - 1] * h, nu, s == 1 ? 1 : 2, last ? 1 : 0));
    }
    const int kl = (kb + ns - 1) % 2;
    if (n + 1 < nsteps) {
      kb = (kl + 1) % 2;
      RC(GF_DISPATCH_N3(m->Niso, launch_stage3b, m, acc, K[kl], u, K[kb], c[ns - 1] * h, 0.0, nu,
                        3, 0));
      have_k0 = true;
    } else {
      RC(GF_DISPATCH_N3(m->Niso, launch_rk_out_box, m, u, acc, K[kl], c[ns - 1] * h));
    }
  }
  return GF_OK;
}

// ---- pure-advection stage on 2-D embedded (sphere) elements ----------------
// One kernel per RK stage computes, per element,
//   us   = uin + bh Kprev            (stage input; Kprev already smoothed)
//   ust  = ust + ch Kprev            (the step_b accumulation of Kprev)
//   K    = imass * ( -nu L us - M (c . grad us) )
// with the advection of GAdvect::apply on an embedded grid (gadvect.ipp:57-99,
// deriv(p,j) = (D1 p) dXidX(0,j) + (D2 p) dXidX(1,j), ggrid.ipp:1179-1200) and
// the Laplacian of embed_prod (ghelmholtz.ipp:205-290).  The advection is
// regrouped as c . grad us = (D1 us) chat_0 + (D2 us) chat_1 with
// chat_i = sum_j dXidX(i,j) c_j: the first stage of a step forms
// w_i = imass M chat_i from the velocity and the metric and stores it, later
// stages read the two w_i instead of three velocity and six metric fields,
//   K = imass (-nu L us) - ((D1 us) w_0 + (D2 us) w_1).
// The regrouping changes rounding only (relative 1e-16 per term).
template <typename T>
struct EmbAdvArgs {
  const T *uin;         // stage input base (state at s = 0, step-start copy after)
  const T *kprev;       // smoothed K of the previous stage, or null
  T *ust;               // state, advanced by ch * kprev when kprev is set
  T *ucopy;             // receives uin when kprev is null (step-start copy)
  int acc_from_uin;     // the accumulation of kprev starts from the stage's uin
                        // value instead of ust (the step's first accumulation)
  double bh, ch, nu;
  const double *c[3];   // velocity (read when wform is set)
  const double *ndx[6]; // dXi_i/dX_j at [i*3+j]
  const double *G[3];
  const double *mass, *imass;
  const T *D;
  double *w[2];         // imass M chat_i: written when wform is set, else read
  int wform;
  T *kout;
  size_t nelems;
};

// T is the precision of the advected field, of its stage fields and of the
// two tensor-product contractions (FP64, or FP32 in the mixed-precision
// stepper); the metric terms (dXi/dX, G, mass, the velocity products w_i)
// stay FP64 and the products with them are formed in FP64.
template <int N, typename T>
__global__ __launch_bounds__(kThreads) void k_adv_stage_embed(EmbAdvArgs<T> a) {
  using S = Shape<2, N>;
  static_assert(S::NODES <= kThreads, "one node per thread");
  __shared__ T su[S::NODES];
  __shared__ T sg[2][S::NODES];
  __shared__ T sD[N * N];
  const size_t e0 = (size_t)blockIdx.x * S::EPB;
  const int nel = (int)min((size_t)S::EPB, a.nelems - e0);
  const size_t base = e0 * S::NN;
  gf::stage_D(sD, a.D, N);
  const int q = threadIdx.x;
  const bool act = q < nel * S::NN;
  const size_t g = base + q;
  if (act) {
    T v = a.uin[g];
    if (a.kprev) {
      const T kp = a.kprev[g];
      const T base = a.acc_from_uin ? v : a.ust[g];
      a.ust[g] = base + (T)a.ch * kp;
      v = v + (T)a.bh * kp;
    } else if (a.ucopy) {
      a.ucopy[g] = v;
    }
    su[q] = v;
  }
  __syncthreads();
  const int el = q / S::NN, r = q - el * S::NN;
  int i, j, k;
  gf::node_ijk<2, N>(r, i, j, k);
  double po = 0.0, im = 0.0;
  if (act) {
    const T *se = su + el * S::NN;
    const double d0 = gf::contract<2, N, false>(se, sD, 0, i, j, k);
    const double d1 = gf::contract<2, N, false>(se, sD, 1, i, j, k);
    im = a.imass[g];
    double w0, w1;
    if (a.wform) {
      const double c0 = a.c[0][g], c1 = a.c[1][g], c2 = a.c[2][g];
      const double mi = im * a.mass[g];
      w0 = mi * ((a.ndx[0][g] * c0 + a.ndx[1][g] * c1) + a.ndx[2][g] * c2);
      w1 = mi * ((a.ndx[3][g] * c0 + a.ndx[4][g] * c1) + a.ndx[5][g] * c2);
      a.w[0][g] = w0;
      a.w[1][g] = w1;
    } else {
      w0 = a.w[0][g];
      w1 = a.w[1][g];
    }
    po = d0 * w0 + d1 * w1;
    const double G00 = a.G[0][g], G01 = a.G[1][g], G11 = a.G[2][g];
    double g0 = 0.0 + d0 * G00;
    g0 += d1 * G01;
    double g1 = 0.0 + d0 * G01;
    g1 += d1 * G11;
    sg[0][q] = (T)g0;
    sg[1][q] = (T)g1;
  }
  __syncthreads();
  if (act) {
    T lt = T(0);
    lt += gf::contract<2, N, true>(&sg[0][el * S::NN], sD, 0, i, j, k);
    lt += gf::contract<2, N, true>(&sg[1][el * S::NN], sD, 1, i, j, k);
    double lap = lt;
    if (a.nu != 1.0) lap *= a.nu;
    a.kout[g] = (T)(im * (-lap) - po);
  }
}

template <typename T>
struct AdvStageLaunch {
  template <int DIM, int N>
  static int run(gf_mesh *m, const EmbAdvArgs<T> &a) {
    if (DIM != 2) return GF_ERR_UNSUPPORTED;
    using S = Shape<2, N>;
    const size_t nb = (m->nelems + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL((k_adv_stage_embed<N, T>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, a);
    GF_LAUNCH_CHECK(m->ctx);
    return GF_OK;
  }
};
template <int DIM, int N>
int launch_adv_stage_embed_d(gf_mesh *m, const EmbAdvArgs<double> &a) {
  return AdvStageLaunch<double>::run<DIM, N>(m, a);
}
template <int DIM, int N>
int launch_adv_stage_embed_f(gf_mesh *m, const EmbAdvArgs<float> &a) {
  return AdvStageLaunch<float>::run<DIM, N>(m, a);
}
int launch_adv_stage_embed_t(gf_mesh *m, const EmbAdvArgs<double> &a) {
  return GF_DISPATCH_N(2, m->Niso, launch_adv_stage_embed_d, m, a);
}
int launch_adv_stage_embed_t(gf_mesh *m, const EmbAdvArgs<float> &a) {
  return GF_DISPATCH_N(2, m->Niso, launch_adv_stage_embed_f, m, a);
}

template <typename T>
__global__ void k_final_input_t(T *__restrict__ u3, const T *__restrict__ uin,
                                const T *__restrict__ k, T bh, T *acc, const T *accsrc, T ch,
                                size_t n) {
  XCD_GRID_LOOP(g, n) {
    T kv = k[g];
    u3[g] = uin[g] + bh * kv;
    if (acc) acc[g] = accsrc[g] + kv * ch;
  }
}
template <typename T>
__global__ void k_axpby_t(T *__restrict__ z, T a, const T *__restrict__ x, T b, const T *y,
                          size_t n) {
  XCD_GRID_LOOP(g, n) z[g] = a * x[g] + b * y[g];
}
int dss_smooth(gf_mesh *m, double *u) { return gf_dss_t(m, GF_DSS_SMOOTH, u); }
int dss_smooth(gf_mesh *m, float *u) { return gf_dss_f32(m, GF_DSS_SMOOTH, u); }

// Pure-advection RK steps on a 2-D embedded isotropic mesh with three velocity
// components (mode 1, nstate 4): the arithmetic of rk_step_general with each
// stage's RHS, stage input and state accumulation fused into one element
// kernel.  Stage sequence (step_b, gexrk_stepper.ipp:251-341):
//   s < ns-1:  K_s = F(uin + b_{s,s-1} h K_{s-1});  Smooth K_s;  ust += c_s h K_s
//   last:      us = uin + b h K_{ns-2};  Smooth us;  K = F(us);  Smooth K;
//              ust += c h K;  Smooth ust
// where the ust accumulation of K_{s-1} is done by stage s's kernel, the
// first one from the step-start copy of uin (equal to ust at that point).
// The Butcher tables of orders 1-4 have only sub-diagonal b, so the
// reference's isum = sum_j b_{sj} h K_j equals b_{s,s-1} h K_{s-1} exactly.
// (Forming the Smooths by gather in the reading kernels instead of the DSS
// passes, from a per-node partner map, measured slower on the sphere: the
// dependent partner loads cost each kernel more than the pass they save.)
// T = float is the mixed-precision form: the advected field and its stage
// fields in FP32, the metric terms in FP64 (k_adv_stage_embed).
template <typename T>
int rk_steps_embed_adv_t(gf_mesh *m, double nu, int order, double dt, T *ust0,
                         const double *const *vel, int nsteps) {
  const size_t n = m->ndof, bytes = n * sizeof(T), dbytes = n * sizeof(double);
  double alpha[4], beta[16], c[4];
  butcher(order, alpha, beta, c);
  const int ns = order;
  const double h = dt;
  T *ucopy = (T *)gf_ws(m, WS_RK_BEGIN + 0, bytes);
  T *us = (T *)gf_ws(m, WS_RK_BEGIN + 3, bytes);
  T *K[2] = {(T *)gf_ws(m, WS_RK_BEGIN + 6, bytes), (T *)gf_ws(m, WS_RK_BEGIN + 9, bytes)};
  double *w0 = (double *)gf_ws(m, WS_RK_BEGIN + 1, dbytes);
  double *w1 = (double *)gf_ws(m, WS_RK_BEGIN + 2, dbytes);
  if (!ucopy || !us || !K[0] || !K[1] || !w0 || !w1)
    return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  EmbAdvArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.nu = nu;
  for (int k = 0; k < 3; k++) a.c[k] = vel[k];
  for (int k = 0; k < 6; k++) a.ndx[k] = m->ndx[(k / 3) * 3 + k % 3];
  for (int k = 0; k < 3; k++) a.G[k] = m->G[k];
  a.mass = m->mass;
  a.imass = m->imass;
  a.D = sizeof(T) == 8 ? (const T *)m->D[0] : (const T *)m->D32;
  a.w[0] = w0;
  a.w[1] = w1;
  a.nelems = m->nelems;
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < ns - 1; s++) {
      a.wform = s == 0;
      a.uin = s == 0 ? ust0 : ucopy;
      a.acc_from_uin = s == 1;
      a.kprev = s == 0 ? nullptr : K[(s - 1) % 2];
      a.ust = ust0;
      a.ucopy = s == 0 ? ucopy : nullptr;
      a.bh = s == 0 ? 0.0 : beta[s * 4 + (s - 1)] * h;
      a.ch = s == 0 ? 0.0 : c[s - 1] * h;
      a.kout = K[s % 2];
      RC(launch_adv_stage_embed_t(m, a));
      RC(dss_smooth(m, K[s % 2]));
    }
    if (ns >= 2) {
      // order 2: the step's first accumulation, from the step-start copy
      hipLaunchKernelGGL(k_final_input_t<T>, dim3(xcd_grid(nblocks(n))), dim3(kThreads), 0,
                         m->ctx->stream, us, ucopy, K[(ns - 2) % 2],
                         (T)(beta[(ns - 1) * 4 + (ns - 2)] * h), ust0, ns == 2 ? ucopy : ust0,
                         (T)(c[ns - 2] * h), n);
      GF_LAUNCH_CHECK(m->ctx);
    } else {
      GF_HIP(m->ctx, hipMemcpyAsync(us, ust0, bytes, hipMemcpyDeviceToDevice, m->ctx->stream));
    }
    RC(dss_smooth(m, us));
    a.uin = us;
    a.acc_from_uin = 0;
    a.kprev = nullptr;
    a.ucopy = nullptr;
    a.wform = ns == 1;
    a.bh = a.ch = 0.0;
    a.kout = K[(ns - 1) % 2];
    RC(launch_adv_stage_embed_t(m, a));
    T *kl = K[(ns - 1) % 2];
    RC(dss_smooth(m, kl));
    hipLaunchKernelGGL(k_axpby_t<T>, dim3(xcd_grid(nblocks(n))), dim3(kThreads), 0,
                       m->ctx->stream, ust0, (T)1, ust0, (T)(c[ns - 1] * h), kl, n);
    GF_LAUNCH_CHECK(m->ctx);
    RC(dss_smooth(m, ust0));
  }
  return GF_OK;
}

int rk_step_embed_adv(gf_mesh *m, double nu, int order, double dt, double *const *ust,
                      int nsteps) {
  return rk_steps_embed_adv_t<double>(m, nu, order, dt, ust[0], ust + 1, nsteps);
}

// General RK step composed from the operators, following step_b literally
// (gexrk_stepper.ipp:251-341).  Used for embedded meshes, pure advection and
// heat modes, and orders without a fused path.
int rk_step_general(gf_mesh *m, int mode, double nu, int order, int nstate, double dt,
                    double *const *ust) {
  size_t n = m->ndof, bytes = n * sizeof(double);
  int nev = mode == 0 ? nstate : 1;
  if (nev > 3) return gf_set_err(m->ctx, GF_ERR_ARG, "at most 3 evolved components");
  double alpha[4], beta[16], c[4];
  butcher(order, alpha, beta, c);
  int ns = order;
  double h = dt;
  double *uin[3], *us[3], *K[3][3], *isum;
  for (int k = 0; k < nev; k++) {
    uin[k] = (double *)gf_ws(m, WS_RK_BEGIN + k, bytes);
    us[k] = (double *)gf_ws(m, WS_RK_BEGIN + 3 + k, bytes);
    for (int s = 0; s < std::max(ns - 1, 1); s++) {
      K[s][k] = (double *)gf_ws(m, WS_RK_BEGIN + 6 + 3 * s + k, bytes);
      if (!K[s][k]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
    }
    if (!uin[k] || !us[k]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
    RC(copy(m, uin[k], ust[k]));
  }
  isum = (double *)gf_ws(m, WS_RK_BEGIN + 15, bytes);
  if (!isum) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  std::vector<const double *> rin(nstate);
  for (int k = nev; k < nstate; k++) rin[k] = ust[k];
  // heat mode evolves only u[0] (gburgers.ipp:295-300); its RHS sees that
  // component alone, so K never needs more than nev outputs.
  const int nrhs = mode == 2 ? nev : nstate;
  for (int s = 0; s < ns - 1; s++) {
    for (int k = 0; k < nev; k++) {
      GF_HIP(m->ctx, hipMemsetAsync(isum, 0, bytes, m->ctx->stream));
      for (int j = 0; j < s; j++) RC(axpby(m, isum, 1.0, isum, beta[s * 4 + j] * h, K[j][k]));
      RC(axpby(m, us[k], 1.0, uin[k], 1.0, isum));
    }
    RC(constrain(m, us, nev));
    for (int k = 0; k < nev; k++) rin[k] = us[k];
    RC(gf_burgers_rhs(m, mode, nu, nrhs, rin.data(), K[s]));
    for (int k = 0; k < nev; k++) {
      RC(gf_dss_t(m, GF_DSS_SMOOTH, K[s][k]));
      RC(axpby(m, ust[k], 1.0, ust[k], c[s] * h, K[s][k]));
    }
    RC(constrain(m, ust, nev));
  }
  for (int k = 0; k < nev; k++) {
    GF_HIP(m->ctx, hipMemsetAsync(isum, 0, bytes, m->ctx->stream));
    for (int j = 0; j < ns - 1; j++) RC(axpby(m, isum, 1.0, isum, beta[(ns - 1) * 4 + j] * h, K[j][k]));
    RC(axpby(m, us[k], 1.0, uin[k], 1.0, isum));
    RC(gf_dss_t(m, GF_DSS_SMOOTH, us[k]));
  }
  RC(constrain(m, us, nev));
  for (int k = 0; k < nev; k++) rin[k] = us[k];
  RC(gf_burgers_rhs(m, mode, nu, nrhs, rin.data(), K[0]));
  for (int k = 0; k < nev; k++) RC(gf_dss_t(m, GF_DSS_SMOOTH, K[0][k]));
  for (int k = 0; k < nev; k++) {
    RC(axpby(m, ust[k], 1.0, ust[k], c[ns - 1] * h, K[0][k]));
    RC(gf_dss_t(m, GF_DSS_SMOOTH, ust[k]));
  }
  RC(constrain(m, ust, nev));
  RC(constrain(m, ust, nev));  // GBurgers::step_exrk (gburgers.ipp:405)
  return GF_OK;
}

int check_finite(gf_mesh *m, double *const *u, int nev) {
  GF_HIP(m->ctx, hipMemsetAsync(m->scal + 62, 0, sizeof(double), m->ctx->stream));
  for (int k = 0; k < nev; k++)
    hipLaunchKernelGGL(k_nonfinite, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream,
                       u[k], m->ndof, m->scal + 62);
  GF_LAUNCH_CHECK(m->ctx);
  double f = 0;
  GF_HIP(m->ctx, hipMemcpyAsync(&f, m->scal + 62, sizeof(double), hipMemcpyDeviceToHost,
                                m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return f != 0.0 ? gf_set_err(m->ctx, GF_ERR_NONFINITE, "Solution not finite") : GF_OK;
}

}  // namespace

extern "C" {

int gf_burgers_step(gf_mesh *m, int mode, double nu, int order, int nstate, double t, double dt,
                    double *const *u, int check) {
  (void)t;
  if (!m || !u || order < 1 || order > 4 || nstate < 1 || nstate > 4) return GF_ERR_ARG;
  for (int k = 0; k < nstate; k++)
    if (!u[k]) return GF_ERR_ARG;
  bool fused = mode == 0 && m->gtype == GF_GE_REGULAR && m->Niso && nstate == m->dim &&
               (m->dim == 2 || m->dconst);
  int rc;
  const bool embed_adv = mode == 1 && m->gtype == GF_GE_2DEMBEDDED && m->Niso && nstate == 4;
  if (fused && m->has_box && order >= 2)
    rc = rk_steps_box(m, nu, order, dt, 1, u);
  else if (fused)
    rc = rk_step_fused(m, nu, order, dt, u);
  else if (embed_adv)
    rc = rk_step_embed_adv(m, nu, order, dt, u, 1);
  else
    rc = rk_step_general(m, mode, nu, order, nstate, dt, u);
  if (rc) return rc;
  if (check) return check_finite(m, u, mode == 0 ? nstate : 1);
  return GF_OK;
}

int gf_burgers_steps(gf_mesh *m, int mode, double nu, int order, int nstate, double t, double dt,
                     int nsteps, double *const *u) {
  if (!m || !u || nsteps < 0) return GF_ERR_ARG;
  if (nsteps > 1 && mode == 0 && m->gtype == GF_GE_REGULAR && m->Niso && m->dim == 3 &&
      m->dconst && m->has_box && nstate == 3 && order >= 2 && order <= 4) {
    for (int k = 0; k < nstate; k++)
      if (!u[k]) return GF_ERR_ARG;
    return rk_steps_box(m, nu, order, dt, nsteps, u);
  }
  if (mode == 1 && m->gtype == GF_GE_2DEMBEDDED && m->Niso && nstate == 4 && order >= 1 &&
      order <= 4) {
    for (int k = 0; k < nstate; k++)
      if (!u[k]) return GF_ERR_ARG;
    return rk_step_embed_adv(m, nu, order, dt, u, nsteps);
  }
  for (int s = 0; s < nsteps; s++) {
    int rc = gf_burgers_step(m, mode, nu, order, nstate, t + s * dt, dt, u, 0);
    if (rc) return rc;
  }
  return GF_OK;
}

int gf_burgers_steps_mixed(gf_mesh *m, int mode, double nu, int order, int nstate, double t,
                           double dt, int nsteps, double *const *u) {
  if (!m || !u || nsteps < 0 || order < 1 || order > 4) return GF_ERR_ARG;
  if (!(mode == 1 && nstate == 4 && m->gtype == GF_GE_2DEMBEDDED && m->Niso && m->D32))
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED,
                      "mixed-precision steps: pure advection (mode 1, nstate 4) on embedded meshes");
  for (int k = 0; k < nstate; k++)
    if (!u[k]) return GF_ERR_ARG;
  (void)t;
  const size_t n = m->ndof;
  float *h32 = (float *)gf_ws(m, WS_RK_BEGIN + 4, n * sizeof(float));
  if (!h32) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, m->ctx->stream, h32, u[0], 1.0,
                     n);
  RC(rk_steps_embed_adv_t<float>(m, nu, order, dt, h32, u + 1, nsteps));
  hipLaunchKernelGGL(k_from_f32, dim3(nblocks(n)), dim3(kThreads), 0, m->ctx->stream, u[0], h32, n);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int gf_helmholtz_diag_inv(gf_mesh *m, double p, double qs, const double *qf, double *dinv) {
  if (!m || !dinv) return GF_ERR_ARG;
  if (!m->Niso) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "Jacobi diagonal: isotropic meshes");
  RC(GF_DISPATCH_N(m->dim, m->Niso, launch_diag, m, p, qf, qs, dinv));
  RC(gf_dss_t(m, GF_DSS_SUM, dinv));
  // a Jacobi preconditioner needs a positive assembled diagonal; a strongly
  // sheared deformed element (whose doubled off-diagonal metric terms enter
  // it) or p = q = 0 can break that
  double *bad = m->scal + 63;
  GF_HIP(m->ctx, hipMemsetAsync(bad, 0, sizeof(double), m->ctx->stream));
  hipLaunchKernelGGL(k_recip_pos, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, dinv,
                     bad, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  double hb = 0.0;
  GF_HIP(m->ctx, hipMemcpyAsync(&hb, bad, sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  if (hb != 0.0)
    return gf_set_err(m->ctx, GF_ERR_PRECOND, "Jacobi: assembled diagonal not positive");
  return GF_OK;
}

}  // extern "C"

// ===========================================================================
// CG driver
// ===========================================================================
namespace {

// Reduce (r.z.imult, norm(r)) at set-up, synchronous; returns both.
template <typename T>
int setup_red(gf_mesh *m, const T *r, const T *z, const T *imult, int normtype, double out[2]) {
  hipLaunchKernelGGL(k_setup_red<T>, dim3(kUpdBlocks), dim3(kThreads), 0, m->ctx->stream, r, z,
                     imult, m->mass, normtype, m->red, m->ndof);
  hipLaunchKernelGGL(k_setup_fin, dim3(1), dim3(kThreads), 0, m->ctx->stream, m->red, kUpdBlocks,
                     normtype, m->scal + 60);
  GF_LAUNCH_CHECK(m->ctx);
  GF_HIP(m->ctx, hipMemcpyAsync(out, m->scal + 60, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return GF_OK;
}

double finish_norm_host(int normtype, double v, gf_mesh *m, double volume) {
  switch (normtype) {
    case GF_NORM_INF: return v;
    case GF_NORM_EUC: return std::sqrt(v / (double)m->ndof);
    case GF_NORM_L2: return std::sqrt(v / volume);
    case GF_NORM_L1: return v / volume;
    default: return std::sqrt(v);
  }
}

template <typename T, int NORM>
int launch_update_box(gf_mesh *m, CGVecs<T> &v, double *part) {
  switch (m->Niso) {
    case 2: return launch_update_box_n<T, NORM, 2>(m, v, part);
    case 3: return launch_update_box_n<T, NORM, 3>(m, v, part);
    case 4: return launch_update_box_n<T, NORM, 4>(m, v, part);
    case 5: return launch_update_box_n<T, NORM, 5>(m, v, part);
    case 6: return launch_update_box_n<T, NORM, 6>(m, v, part);
    case 7: return launch_update_box_n<T, NORM, 7>(m, v, part);
    case 8: return launch_update_box_n<T, NORM, 8>(m, v, part);
    case 9: return launch_update_box_n<T, NORM, 9>(m, v, part);
    case 10: return launch_update_box_n<T, NORM, 10>(m, v, part);
    default: return GF_ERR_UNSUPPORTED;
  }
}

template <typename T, int NORM>
void launch_update(gf_mesh *m, CGVecs<T> &v, double *part, bool mask_in_q) {
  const T *mk = (m->has_mask && !mask_in_q) ? v.mask : nullptr;
  hipLaunchKernelGGL((k_cg_update<T, NORM>), dim3(kUpdBlocks), dim3(kThreads), 0, m->ctx->stream,
                     v.x, v.w, v.r, v.q, mk, v.dinv, v.imult, m->mass, m->scal, part, m->ndof);
}

// Runs CG iterations from the state prepared in scal (rho, rtol, first=1,
// iter=0, done) until done.  r holds the masked assembled residual, x the
// iterate.  Returns iterations performed.
template <typename T>
int cg_loop(gf_mesh *m, CGVecs<T> &v, int normtype, double *d_resid, int *iters) {
  hipStream_t s = m->ctx->stream;
  const bool col = m->dim == 3 && m->dconst;
  // Structured box: the update assembles q by gather (mask already in q).
  const bool gather = col && m->has_box;
  int nbA = 0;
  if (col) {
    nbA = col_blocks(m);
  } else {
    int NN = (int)m->nn;
    int epb = NN >= kThreads ? 1 : kThreads / NN;
    nbA = (int)((m->nelems + epb - 1) / epb);
  }
  double *partA = m->red;
  double *partU = m->red + (size_t)nbA + 64;
  CGApplyArgs a;
  a.r = v.r;
  a.dinv = v.dinv;
  a.w = v.w;
  a.q = v.q;
  a.D = v.D;
  a.W = v.W;
  a.edx = v.edx;
  a.ejac = v.ejac;
  a.qf = v.qf;
  a.qs = v.qs;
  for (int c = 0; c < 6; c++) a.G[c] = v.G[c];
  a.mass = v.mass;
  a.p = v.p;
  a.scal = m->scal;
  a.part = partA;
  a.nelems = m->nelems;
  CGApply3Args a3;
  a3.seg_eoff = nullptr;
  a3.nseg = 0;
  a3.sscal = nullptr;
  a3.r = v.r;
  a3.dinv = v.dinv;
  a3.dtab = v.dtab;
  a3.deq = v.deq;
  a3.w = v.w;
  a3.q = v.q;
  a3.D = v.D;
  for (int c = 0; c < 3; c++) {
    a3.DR[c] = sizeof(T) == 8 ? (const void *)m->pDR[c] : (const void *)m->pDR32[c];
    a3.DC[c] = sizeof(T) == 8 ? (const void *)m->pDC[c] : (const void *)m->pDC32[c];
  }
  a3.W = v.W;
  a3.edx = v.edx;
  a3.ejac = v.ejac;
  a3.qf = v.qf;
  a3.qs = v.qs;
  // the mask is formed from the box faces when it is known to match them
  a3.tmask = (m->has_mask && v.mask && m->has_box_mask) ? 1 : 0;
  a3.topo = m->box;
  a3.mask = (m->has_mask && !a3.tmask) ? v.mask : nullptr;
  a3.p = v.p;
  a3.scal = m->scal;
  a3.part = partA;
  a3.nelems = m->nelems;
  // the host mirror starts from the scalars prepared by the caller (a
  // solve that is already done queues nothing that would update it)
  volatile double *hm = m->hscal;
  GF_HIP(m->ctx, hipMemcpyAsync(m->hscal, m->scal, S_COUNT * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  if (hm[S_DONE] != 0.0) {
    *iters = (int)m->hscal[S_ITER];
    return GF_OK;
  }
  int cur = 0, nq = 0;
  for (;;) {
    for (int b = 0; b < kBatch; b++) {
      int rc;
      if (col)
        rc = sizeof(T) == 8 ? GF_DISPATCH_N3(m->Niso, cgapply3c_d, m, a3)
                            : GF_DISPATCH_N3(m->Niso, cgapply3c_f, m, a3);
      else
        rc = sizeof(T) == 8 ? GF_DISPATCH_N(m->dim, m->Niso, launch_cg_apply_d, m, a)
                            : GF_DISPATCH_N(m->dim, m->Niso, launch_cg_apply_f, m, a);
      if (rc) return rc;
      if (m->ngroups && !gather)
        hipLaunchKernelGGL(k_dss_cg<T>, dim3(nblocks((size_t)m->ngroups)), dim3(kThreads), 0, s,
                           v.q, m->goff, m->gmem, m->ngroups, (const double *)m->scal);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kRedThreads), 0, s, m->scal, partA, nbA);
      if (gather) {
        switch (normtype) {
          case GF_NORM_INF: rc = launch_update_box<T, GF_NORM_INF>(m, v, partU); break;
          case GF_NORM_EUC: rc = launch_update_box<T, GF_NORM_EUC>(m, v, partU); break;
          case GF_NORM_L2: rc = launch_update_box<T, GF_NORM_L2>(m, v, partU); break;
          case GF_NORM_L1: rc = launch_update_box<T, GF_NORM_L1>(m, v, partU); break;
          default: rc = launch_update_box<T, GF_NORM_ASSEMBLED_L2>(m, v, partU); break;
        }
        if (rc) return rc;
      } else {
        switch (normtype) {
          case GF_NORM_INF: launch_update<T, GF_NORM_INF>(m, v, partU, col); break;
          case GF_NORM_EUC: launch_update<T, GF_NORM_EUC>(m, v, partU, col); break;
          case GF_NORM_L2: launch_update<T, GF_NORM_L2>(m, v, partU, col); break;
          case GF_NORM_L1: launch_update<T, GF_NORM_L1>(m, v, partU, col); break;
          default: launch_update<T, GF_NORM_ASSEMBLED_L2>(m, v, partU, col); break;
        }
      }
      hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(kThreads), 0, s, m->scal, partU, kUpdBlocks,
                         normtype, d_resid, m->hscal_dev);
    }
    GF_LAUNCH_CHECK(m->ctx);
    // One batch of look-ahead: batch b is queued before the host waits for
    // batch b-1, whose last k_cg_beta has written the stopping flag into
    // the mapped host mirror.  Iterations queued after convergence return
    // at once (S_DONE).
    GF_HIP(m->ctx, hipEventRecord(m->cg_ev[cur], s));
    if (nq++ > 0) {
      GF_HIP(m->ctx, hipEventSynchronize(m->cg_ev[cur ^ 1]));
      if (hm[S_DONE] != 0.0) break;
    }
    cur ^= 1;
  }
  // the final scalars, after everything queued has run
  GF_HIP(m->ctx, hipMemcpyAsync(m->hscal, m->scal, S_COUNT * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  *iters = (int)m->hscal[S_ITER];
  return GF_OK;
}

}  // namespace

namespace {

// Device vectors of one CG solve: FP64 residual, direction, operator output,
// Jacobi diagonal inverse and residual history; the FP32 recurrence vectors
// of the mixed solver.
struct CGWork {
  double *r = nullptr, *w = nullptr, *q = nullptr, *dinv = nullptr, *resid = nullptr;
  float *r32 = nullptr, *w32 = nullptr, *q32 = nullptr, *x32 = nullptr, *dinv32 = nullptr,
        *qf32 = nullptr;
};

int cg_check_args(gf_mesh *m, const gf_cg_params *prm, const double *b, double *x,
                  gf_cg_result *res) {
  if (!m || !prm || !b || !x || !res) return GF_ERR_ARG;
  memset(res, 0, sizeof(*res));
  if (!m->Niso) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "CG: isotropic meshes");
  if (prm->maxit < 1) return gf_set_err(m->ctx, GF_ERR_BADITERNUM, "maxit < 1");
  if (prm->normtype < GF_NORM_INF || prm->normtype > GF_NORM_ASSEMBLED_L2) return GF_ERR_ARG;
  if (prm->p_field) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "CG: p field not supported");
  return GF_OK;
}

// ---- CG on a batched mesh (members set by gf_mesh_set_segments) -----------
__global__ void k_fill_d(double *__restrict__ o, double v, size_t n) {
  GRID_LOOP(g, n) o[g] = v;
}

// Per-member (rho, norm sum) of r (and z): out[2s], out[2s+1]; synchronous.
template <typename T>
int seg_red(gf_mesh *m, const T *r, const T *z, const T *imult, int normtype,
            std::vector<double> &out) {
  const int ns = m->nseg, nbU = m->h_seg_boff[ns];
  double *dout = m->red + 2 * (size_t)nbU + 64;
  hipLaunchKernelGGL(k_setup_red_seg<T>, dim3(nbU), dim3(kThreads), 0, m->ctx->stream, r, z, imult,
                     m->mass, normtype, m->seg_eoff, m->seg_boff, ns, m->nn, m->red);
  hipLaunchKernelGGL(k_setup_fin_seg, dim3(1), dim3(1024), 0, m->ctx->stream, m->red, m->seg_boff,
                     ns, normtype, dout);
  GF_LAUNCH_CHECK(m->ctx);
  out.resize(2 * (size_t)ns);
  GF_HIP(m->ctx, hipMemcpyAsync(out.data(), dout, 2 * ns * sizeof(double), hipMemcpyDeviceToHost,
                                m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return GF_OK;
}

// CG iterations of every member of a batched 3-D mesh (pencil apply, CSR
// assembly) until all members are done; per-member scalars in m->sscal,
// copied to m->hsscal at the end.
template <typename T>
int cg_loop_seg(gf_mesh *m, CGVecs<T> &v, int normtype) {
  hipStream_t s = m->ctx->stream;
  const int ns = m->nseg, nbU = m->h_seg_boff[ns];
  const int nbA = col_blocks(m);
  double *partA = m->red;  // one partial per element
  double *partU = m->red + std::max((size_t)nbA, m->nelems) + 64;
  CGApply3Args a3;
  a3.r = v.r;
  a3.dinv = v.dinv;
  a3.dtab = nullptr;
  a3.deq = nullptr;
  a3.w = v.w;
  a3.q = v.q;
  a3.D = v.D;
  for (int c = 0; c < 3; c++) {
    a3.DR[c] = sizeof(T) == 8 ? (const void *)m->pDR[c] : (const void *)m->pDR32[c];
    a3.DC[c] = sizeof(T) == 8 ? (const void *)m->pDC[c] : (const void *)m->pDC32[c];
  }
  a3.W = v.W;
  a3.edx = v.edx;
  a3.ejac = v.ejac;
  a3.qf = v.qf;
  a3.qs = v.qs;
  a3.tmask = 0;
  a3.topo = m->box;
  a3.mask = m->has_mask ? v.mask : nullptr;
  a3.p = v.p;
  a3.scal = m->scal;
  a3.part = partA;
  a3.nelems = m->nelems;
  a3.seg_eoff = m->seg_eoff;
  a3.nseg = ns;
  a3.sscal = m->sscal;
  volatile double *hm = m->hscal;
  GF_HIP(m->ctx, hipMemcpyAsync(m->hscal, m->scal, S_COUNT * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  int cur = 0, nq = 0;
  while (hm[S_DONE] == 0.0) {
    for (int bb = 0; bb < kBatch; bb++) {
      int rc = sizeof(T) == 8 ? GF_DISPATCH_N3(m->Niso, cgapply3c_d, m, a3)
                              : GF_DISPATCH_N3(m->Niso, cgapply3c_f, m, a3);
      if (rc) return rc;
      if (m->ngroups)
        hipLaunchKernelGGL(k_dss_cg<T>, dim3(nblocks((size_t)m->ngroups)), dim3(kThreads), 0, s,
                           v.q, m->goff, m->gmem, m->ngroups, (const double *)m->scal);
      hipLaunchKernelGGL(k_cg_alpha_seg, dim3(1), dim3(1024), 0, s, m->sscal, partA, m->seg_eoff,
                         ns, 1);
      const T *mk = m->has_mask ? v.mask : nullptr;
      switch (normtype) {
#define GF_SEG_UPD(NT)                                                                        \
  hipLaunchKernelGGL((k_cg_update_seg<T, NT>), dim3(nbU), dim3(kThreads), 0, s, v.x, v.w, v.r, \
                     v.q, mk, v.dinv, v.imult, m->mass, (const double *)m->sscal, m->seg_eoff,  \
                     m->seg_boff, ns, m->nn, partU);                                           \
  break;
        case GF_NORM_INF: GF_SEG_UPD(GF_NORM_INF)
        case GF_NORM_EUC: GF_SEG_UPD(GF_NORM_EUC)
        case GF_NORM_L2: GF_SEG_UPD(GF_NORM_L2)
        case GF_NORM_L1: GF_SEG_UPD(GF_NORM_L1)
        default: GF_SEG_UPD(GF_NORM_ASSEMBLED_L2)
#undef GF_SEG_UPD
      }
      hipLaunchKernelGGL(k_cg_beta_seg, dim3(1), dim3(1024), 0, s, m->scal, m->sscal, partU,
                         m->seg_boff, ns, normtype, m->hscal_dev);
    }
    GF_LAUNCH_CHECK(m->ctx);
    GF_HIP(m->ctx, hipEventRecord(m->cg_ev[cur], s));
    if (nq++ > 0) {
      GF_HIP(m->ctx, hipEventSynchronize(m->cg_ev[cur ^ 1]));
      if (hm[S_DONE] != 0.0) break;
    }
    cur ^= 1;
  }
  GF_HIP(m->ctx, hipMemcpyAsync(m->hsscal, m->sscal, (size_t)ns * kSegScal * sizeof(double),
                                hipMemcpyDeviceToHost, s));
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  return GF_OK;
}

// gf_cg_solve on a batched mesh: the GCG recurrence (FP64) or the
// mixed-precision solver of cg_solve_impl, with every member's own scalars,
// norm of b, stopping test and reliable updates.
int cg_solve_seg(gf_mesh *m, const gf_cg_params *prm, const double *b, const double *xb,
                 double *x, gf_cg_result *res, const CGWork &wk) {
  if (!(m->dim == 3 && m->dconst && !m->has_box))
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED,
                      "batched CG: 3-D regular meshes with the standard D, CSR assembly");
  hipStream_t s = m->ctx->stream;
  const int ns = m->nseg;
  const size_t n = m->ndof;
  const bool bbv = xb != nullptr;
  const int normtype = prm->normtype;
  double *r = wk.r, *w = wk.w, *q = wk.q, *dinv = wk.dinv;
  const double qs = prm->q_field ? 0.0 : prm->q_scalar;
  std::vector<double> o, vol(ns), nseg_n(ns), rn0(ns), rtol(ns);
  for (int k = 0; k < ns; k++)
    nseg_n[k] = (double)((m->h_seg_eoff[k + 1] - m->h_seg_eoff[k]) * (int64_t)m->nn);
  // member volumes: sum of the mass (the L1 norm of ones)
  hipLaunchKernelGGL(k_fill_d, dim3(nblocks(n)), dim3(kThreads), 0, s, w, 1.0, n);
  RC(seg_red<double>(m, w, nullptr, m->imult, GF_NORM_L1, o));
  for (int k = 0; k < ns; k++) vol[k] = o[2 * k + 1];
  auto fin = [&](int k, double v) {
    switch (normtype) {
      case GF_NORM_INF: return v;
      case GF_NORM_EUC: return std::sqrt(v / nseg_n[k]);
      case GF_NORM_L2: return std::sqrt(v / vol[k]);
      case GF_NORM_L1: return v / vol[k];
      default: return std::sqrt(v);
    }
  };
  if (prm->precond) RC(gf_helmholtz_diag_inv(m, prm->p_scalar, qs, prm->q_field, dinv));
  if (bbv) {
    RC(gf_mul_t(m, x, m->mask, x));
    RC(gf_axpy_t(m, 1.0, xb, x));
  }
  if (normtype == GF_NORM_ASSEMBLED_L2) {
    RC(copy(m, q, b));
    RC(gf_dss_t(m, GF_DSS_SUM, q));
    if (bbv) RC(gf_mul_t(m, q, m->mask, q));
    RC(seg_red<double>(m, q, nullptr, m->imult, normtype, o));
    for (int k = 0; k < ns; k++) {
      rn0[k] = fin(k, o[2 * k + 1]);
      rtol[k] = prm->tol * rn0[k];
    }
  } else {
    RC(seg_red<double>(m, b, nullptr, m->imult, normtype, o));
    for (int k = 0; k < ns; k++) {
      rn0[k] = fin(k, o[2 * k + 1]);
      rtol[k] = rn0[k] < 1.0 ? prm->tol : prm->tol * rn0[k];
    }
  }
  auto residual = [&](double *rr) -> int {
    RC(gf_helm_t(m, prm->p_scalar, nullptr, prm->q_field, qs, x, q));
    hipLaunchKernelGGL(k_residual, dim3(nblocks(n)), dim3(kThreads), 0, s, rr, b, q, n);
    RC(gf_dss_t(m, GF_DSS_SUM, rr));
    if (bbv) RC(gf_mul_t(m, rr, m->mask, rr));
    return GF_OK;
  };
  std::vector<double> hs((size_t)ns * kSegScal, 0.0);
  auto upload = [&]() -> int {
    bool all = true;
    for (int k = 0; k < ns; k++) all = all && hs[(size_t)k * kSegScal + S_DONE] != 0.0;
    double g[S_COUNT] = {0};
    g[S_DONE] = all ? 1.0 : 0.0;
    GF_HIP(m->ctx, hipMemcpyAsync(m->sscal, hs.data(), hs.size() * sizeof(double),
                                  hipMemcpyHostToDevice, s));
    GF_HIP(m->ctx, hipMemcpyAsync(m->scal, g, sizeof(g), hipMemcpyHostToDevice, s));
    GF_HIP(m->ctx, hipStreamSynchronize(s));
    return GF_OK;
  };
  m->seg_iters.assign(ns, 0);
  m->seg_rnorm.assign(ns, 0.0);
  m->seg_rnorm0 = rn0;
  int status = GF_OK, outer = 0;
  if (prm->precision == 0) {
    RC(residual(r));
    double *z = q;
    if (dinv) RC(gf_mul_t(m, dinv, r, z));
    RC(seg_red<double>(m, r, dinv ? z : r, m->imult, normtype, o));
    for (int k = 0; k < ns; k++) {
      double *h = &hs[(size_t)k * kSegScal];
      h[S_RHO] = o[2 * k];
      h[S_RTOL] = rtol[k];
      h[S_RNORM] = 10.0 * rtol[k];
      h[S_MAXIT] = prm->maxit;
      h[S_DONE] = (0 < prm->maxit && h[S_RNORM] > rtol[k]) ? 0.0 : 1.0;
      h[S_FIRST] = 1.0;
      h[S_VOL] = vol[k];
      h[S_N] = nseg_n[k];
    }
    RC(upload());
    CGVecs<double> v{};
    v.x = x;
    v.r = r;
    v.w = w;
    v.q = q;
    v.dinv = dinv;
    v.mask = bbv ? m->mask : nullptr;
    v.imult = m->imult;
    v.qf = prm->q_field;
    v.qs = qs;
    v.D = m->D[0];
    v.W = m->W[0];
    v.edx = m->edx;
    v.ejac = m->ejac;
    for (int c = 0; c < 6; c++) v.G[c] = m->G[c];
    v.mass = m->mass;
    v.p = prm->p_scalar;
    RC(cg_loop_seg<double>(m, v, normtype));
    for (int k = 0; k < ns; k++) {
      const double *h = &m->hsscal[(size_t)k * kSegScal];
      m->seg_iters[k] = (int)h[S_ITER];
      m->seg_rnorm[k] = h[S_RNORM];
      if (m->seg_iters[k] >= prm->maxit && h[S_RNORM] > prm->tol) status = GF_ERR_NOCONVERGE;
    }
  } else {
    float *r32 = wk.r32, *w32 = wk.w32, *q32 = wk.q32, *x32 = wk.x32;
    float *dinv32 = dinv ? wk.dinv32 : nullptr;
    float *qf32 = prm->q_field ? wk.qf32 : nullptr;
    if (dinv) hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, dinv32, dinv, 1.0, n);
    if (qf32) hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, qf32, prm->q_field, 1.0, n);
    const double delta = prm->inner_tol > 0 ? prm->inner_tol : 0.01;
    CGVecs<float> v{};
    v.x = x32;
    v.r = r32;
    v.w = w32;
    v.q = q32;
    v.dinv = dinv32;
    v.mask = bbv ? m->mask32 : nullptr;
    v.imult = m->imult32;
    v.qf = qf32;
    v.qs = qs;
    v.D = m->D32;
    v.W = m->W32;
    v.edx = m->edx32;
    v.ejac = m->ejac32;
    for (int c = 0; c < 6; c++) v.G[c] = nullptr;
    v.mass = nullptr;
    v.p = prm->p_scalar;
    std::vector<int> first(ns, 1), conv(ns, 0);
    std::vector<double> rhom(ns, 0.0);
    for (outer = 0;; outer++) {
      if (outer > 0)
        hipLaunchKernelGGL(k_add_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, x, x32, 1.0, n);
      RC(residual(r));
      RC(seg_red<double>(m, r, nullptr, m->imult, normtype, o));
      bool all = true;
      for (int k = 0; k < ns; k++) {
        m->seg_rnorm[k] = fin(k, o[2 * k + 1]);
        conv[k] = m->seg_rnorm[k] <= rtol[k] || m->seg_iters[k] >= prm->maxit;
        all = all && conv[k];
      }
      if (all || outer >= 10000) break;
      hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, r32, r, 1.0, n);
      hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, x32, 0.0f, n);
      float *z32 = q32;
      if (dinv32) hipLaunchKernelGGL(k_mul_t<float>, dim3(nblocks(n)), dim3(kThreads), 0, s, dinv32, r32, z32, n);
      RC(seg_red<float>(m, r32, dinv32 ? z32 : r32, m->imult32, normtype, o));
      for (int k = 0; k < ns; k++) {
        double *h = &hs[(size_t)k * kSegScal];
        std::fill(h, h + kSegScal, 0.0);
        if (conv[k]) {
          h[S_DONE] = 1.0;
          continue;
        }
        const double rn32 = fin(k, o[2 * k + 1]);
        h[S_RHO] = o[2 * k];
        h[S_BETA] = first[k] ? 0.0 : o[2 * k] / rhom[k];
        h[S_RHOM] = first[k] ? 0.0 : rhom[k];
        h[S_RTOL] = std::max(delta * rn32, 0.5 * rtol[k]);
        h[S_RNORM] = rn32;
        h[S_MAXIT] = std::max(1, prm->maxit - m->seg_iters[k]);
        h[S_FIRST] = first[k] ? 1.0 : 0.0;
        h[S_VOL] = vol[k];
        h[S_N] = nseg_n[k];
      }
      RC(upload());
      RC(cg_loop_seg<float>(m, v, normtype));
      for (int k = 0; k < ns; k++) {
        if (conv[k]) continue;
        const double *h = &m->hsscal[(size_t)k * kSegScal];
        m->seg_iters[k] += (int)h[S_ITER];
        rhom[k] = h[S_RHOM];
        first[k] = 0;
      }
    }
    for (int k = 0; k < ns; k++)
      if (m->seg_rnorm[k] > rtol[k]) status = GF_ERR_NOCONVERGE;
  }
  if (bbv) {
    RC(gf_mul_t(m, x, m->mask, x));
    RC(gf_axpy_t(m, 1.0, xb, x));
  }
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  res->status = status;
  res->iters = *std::max_element(m->seg_iters.begin(), m->seg_iters.end());
  res->outer = outer;
  res->rnorm = *std::max_element(m->seg_rnorm.begin(), m->seg_rnorm.end());
  res->rnorm0 = *std::max_element(rn0.begin(), rn0.end());
  return status;
}

// Jacobi table of a structured box: dtab = the Jacobi diagonal inverse of
// element dref (FP64 and FP32) and deq[e] = whether element e's values equal
// it bit for bit (uniform elements with complete groups do).
__global__ __launch_bounds__(kThreads) void k_dinv_table(const double *__restrict__ dinv,
                                                         size_t nn, size_t dref, size_t nelems,
                                                         double *__restrict__ dtab,
                                                         float *__restrict__ dtab32,
                                                         unsigned char *__restrict__ deq) {
  const double *ref = dinv + dref * nn;
  for (size_t e = blockIdx.x; e < nelems; e += gridDim.x) {
    int diff = 0;
    for (size_t q = threadIdx.x; q < nn; q += blockDim.x) diff |= dinv[e * nn + q] != ref[q];
    const int any = __syncthreads_or(diff);
    if (threadIdx.x == 0) deq[e] = any ? 0 : 1;
  }
  if (blockIdx.x == 0)
    for (size_t q = threadIdx.x; q < nn; q += blockDim.x) {
      dtab[q] = ref[q];
      dtab32[q] = (float)ref[q];
    }
}

int cg_solve_impl(gf_mesh *m, const gf_cg_params *prm, const double *b, const double *xb,
                  double *x, gf_cg_result *res, const CGWork &wk) {
  hipStream_t s = m->ctx->stream;
  size_t n = m->ndof;
  bool bbv = xb != nullptr;
  const double *mask = bbv ? m->mask : nullptr;
  double *r = wk.r, *w = wk.w, *q = wk.q, *dinv = wk.dinv, *d_resid = wk.resid;
  double volume = 0.0;
  {
    double *one = w;
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(kThreads), 0, s, one, 0.0, b, 0.0, b, n);
    // integrate(1) = sum of mass
    RC(gf_dot_t(m, m->mass, nullptr, n, &volume));
  }
  // mass term: the q field, else the scalar q (GHelmholtz with a one-entry q,
  // ghelmholtz.ipp:184-186, 370-373)
  const double qs = prm->q_field ? 0.0 : prm->q_scalar;
  if (prm->precond) RC(gf_helmholtz_diag_inv(m, prm->p_scalar, qs, prm->q_field, dinv));
  const bool dtabv = prm->precond && m->deq;
  if (dtabv) {
    hipLaunchKernelGGL(k_dinv_table, dim3((unsigned)std::min<size_t>(m->nelems, 65536)),
                       dim3(kThreads), 0, s, dinv, m->nn, m->dref, m->nelems, m->dtab, m->dtab32,
                       m->deq);
    GF_LAUNCH_CHECK(m->ctx);
  }
  if (bbv) {  // x = x*mask + xb (gcg.ipp:371-375)
    RC(gf_mul_t(m, x, m->mask, x));
    RC(gf_axpy_t(m, 1.0, xb, x));
  }
  // stopping reference: norm of b (compute_norm(b), gcg.ipp:290-291)
  double red2[2];
  double rnorm0, rtol;
  int normtype = prm->normtype;
  if (normtype == GF_NORM_ASSEMBLED_L2) {
    RC(copy(m, q, b));
    RC(gf_dss_t(m, GF_DSS_SUM, q));
    if (bbv) RC(gf_mul_t(m, q, m->mask, q));
    RC(setup_red<double>(m, q, nullptr, m->imult, normtype, red2));
    rnorm0 = finish_norm_host(normtype, red2[1], m, volume);
    rtol = prm->tol * rnorm0;
  } else {
    RC(setup_red<double>(m, b, nullptr, m->imult, normtype, red2));
    rnorm0 = finish_norm_host(normtype, red2[1], m, volume);
    rtol = rnorm0 < 1.0 ? prm->tol : prm->tol * rnorm0;
  }
  res->rnorm0 = rnorm0;
  int total_iters = 0, outer = 0, status = GF_OK;
  double rn_final = 0.0;
  // r = b - A x ; DSS ; mask
  auto residual = [&](double *rr) -> int {
    RC(gf_helm_t(m, prm->p_scalar, nullptr, prm->q_field, qs, x, q));
    hipLaunchKernelGGL(k_residual, dim3(nblocks(n)), dim3(kThreads), 0, s, rr, b, q, n);
    RC(gf_dss_t(m, GF_DSS_SUM, rr));
    if (bbv) RC(gf_mul_t(m, rr, m->mask, rr));
    return GF_OK;
  };
  if (prm->precision == 0) {
    RC(residual(r));
    // z = P^-1 r ; w = z (set inside the first apply) ; rho = r.z.imult
    double *z = q;
    if (dinv) RC(gf_mul_t(m, dinv, r, z));
    RC(setup_red<double>(m, r, dinv ? z : r, m->imult, normtype, red2));
    double hs[12] = {0};
    hs[S_RHO] = red2[0];
    hs[S_RTOL] = rtol;
    hs[S_RNORM] = 10.0 * rtol;
    hs[S_ITER] = 0;
    hs[S_MAXIT] = prm->maxit;
    hs[S_DONE] = (0 < prm->maxit && hs[S_RNORM] > rtol) ? 0.0 : 1.0;
    hs[S_FIRST] = 1.0;
    hs[S_VOL] = volume;
    hs[S_N] = (double)n;
    GF_HIP(m->ctx, hipMemcpyAsync(m->scal, hs, 12 * sizeof(double), hipMemcpyHostToDevice, s));
    CGVecs<double> v{};
    v.x = x;
    v.r = r;
    v.w = w;
    v.q = q;
    v.dinv = dinv;
    v.dtab = dtabv ? m->dtab : nullptr;
    v.deq = dtabv ? m->deq : nullptr;
    v.mask = mask;
    v.imult = m->imult;
    v.qf = prm->q_field;
    v.qs = qs;
    v.D = m->D[0];
    v.W = m->W[0];
    v.edx = m->edx;
    v.ejac = m->ejac;
    for (int c = 0; c < 6; c++) v.G[c] = m->G[c];
    v.mass = m->mass;
    v.p = prm->p_scalar;
    int it = 0;
    RC(cg_loop<double>(m, v, normtype, d_resid, &it));
    total_iters = it;
    rn_final = m->hscal[S_RNORM];
    if (prm->resid && it > 0)
      GF_HIP(m->ctx, hipMemcpy(prm->resid, d_resid, it * sizeof(double), hipMemcpyDeviceToHost));
    if (it >= prm->maxit && rn_final > prm->tol) status = GF_ERR_NOCONVERGE;
  } else {
    // Mixed precision: the CG recurrences (apply, direction, residual and
    // iterate increments) run in FP32 with FP64 reductions; whenever the
    // recurrence residual has dropped by the factor `delta` since the last
    // reliable update, the FP32 increment is added to the FP64 iterate and
    // the residual is replaced by the true FP64 residual b - A x, keeping the
    // search direction (reliable updates, van der Vorst & Ye, SIAM J. Sci.
    // Comput. 22 (2000) 1663-1683).  Convergence is judged on the FP64
    // residual only.
    float *r32 = wk.r32, *w32 = wk.w32, *q32 = wk.q32, *x32 = wk.x32;
    float *dinv32 = dinv ? wk.dinv32 : nullptr;
    float *qf32 = prm->q_field ? wk.qf32 : nullptr;
    if (dinv) hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, dinv32, dinv, 1.0, n);
    if (qf32) hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, qf32, prm->q_field, 1.0, n);
    const double delta = prm->inner_tol > 0 ? prm->inner_tol : 0.01;
    CGVecs<float> v{};
    v.x = x32;
    v.r = r32;
    v.w = w32;
    v.q = q32;
    v.dinv = dinv32;
    v.dtab = dtabv ? m->dtab32 : nullptr;
    v.deq = dtabv ? m->deq : nullptr;
    v.mask = bbv ? m->mask32 : nullptr;
    v.imult = m->imult32;
    v.qf = qf32;
    v.qs = qs;
    v.D = m->D32;
    v.W = m->W32;
    v.edx = m->edx32;
    v.ejac = m->ejac32;
    // FP32 copies of the node metrics and the lumped mass (made at mesh
    // creation for embedded and deformed meshes)
    for (int c = 0; c < 6; c++) v.G[c] = m->G32[c];
    v.mass = m->gtype != GF_GE_REGULAR ? m->mass32 : nullptr;
    v.p = prm->p_scalar;
    int hist = 0;
    bool first = true;
    for (outer = 0;; outer++) {
      // reliable update: x += x32 ; r = mask DSS(b - A x) in FP64
      if (!first)
        hipLaunchKernelGGL(k_add_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, x, x32, 1.0, n);
      RC(residual(r));
      RC(setup_red<double>(m, r, nullptr, m->imult, normtype, red2));
      const double rn = finish_norm_host(normtype, red2[1], m, volume);
      rn_final = rn;
      if (prm->resid && hist < prm->maxit) prm->resid[hist++] = rn;
      if (rn <= rtol || total_iters >= prm->maxit || outer >= 10000) break;
      hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, r32, r, 1.0, n);
      hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, x32, 0.0f, n);
      float *z32 = q32;
      if (dinv32) hipLaunchKernelGGL(k_mul_t<float>, dim3(nblocks(n)), dim3(kThreads), 0, s, dinv32, r32, z32, n);
      RC(setup_red<float>(m, r32, dinv32 ? z32 : r32, m->imult32, normtype, red2));
      const double rn32 = finish_norm_host(normtype, red2[1], m, volume);
      double hs[S_COUNT] = {0};
      hs[S_RHO] = red2[0];
      hs[S_BETA] = first ? 0.0 : red2[0] / m->hscal[S_RHOM];
      hs[S_RHOM] = first ? 0.0 : m->hscal[S_RHOM];
      hs[S_RTOL] = std::max(delta * rn32, 0.5 * rtol);
      hs[S_RNORM] = rn32;
      hs[S_ITER] = 0;
      hs[S_MAXIT] = std::max(1, prm->maxit - total_iters);
      hs[S_DONE] = 0.0;
      hs[S_FIRST] = first ? 1.0 : 0.0;
      hs[S_VOL] = volume;
      hs[S_N] = (double)n;
      GF_HIP(m->ctx, hipMemcpyAsync(m->scal, hs, sizeof(hs), hipMemcpyHostToDevice, s));
      int it = 0;
      RC(cg_loop<float>(m, v, normtype, (double *)nullptr, &it));
      total_iters += it;
      first = false;
    }
    if (rn_final > rtol) status = GF_ERR_NOCONVERGE;
  }
  if (bbv) {  // x *= mask ; x += xb (gcg.ipp:337, 378)
    RC(gf_mul_t(m, x, m->mask, x));
    RC(gf_axpy_t(m, 1.0, xb, x));
  }
  GF_HIP(m->ctx, hipStreamSynchronize(s));
  res->status = status;
  res->iters = total_iters;
  res->outer = outer;
  res->rnorm = rn_final;
  return status;
}

}  // namespace

__global__ void k_groups_in_segments(const int *__restrict__ goff, const int *__restrict__ gmem,
                                     int ngroups, const int64_t *__restrict__ eoff, int nseg,
                                     size_t nn, double *bad) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += gridDim.x * blockDim.x) {
    const int a = gmem[goff[g]], z = gmem[goff[g + 1] - 1];
    if (seg_of(eoff, nseg, (size_t)a / nn) != seg_of(eoff, nseg, (size_t)z / nn)) *bad = 1.0;
  }
}

extern "C" int gf_mesh_set_segments(gf_mesh *m, int nseg, const int64_t *eoff) {
  if (!m || nseg < 0 || (nseg > 0 && !eoff)) return GF_ERR_ARG;
  (void)hipFree(m->seg_eoff);
  (void)hipFree(m->seg_boff);
  (void)hipFree(m->sscal);
  if (m->hsscal) (void)hipHostFree(m->hsscal);
  m->seg_eoff = nullptr;
  m->seg_boff = nullptr;
  m->sscal = m->hsscal = m->hsscal_dev = nullptr;
  m->nseg = 0;
  m->h_seg_eoff.clear();
  m->h_seg_boff.clear();
  if (nseg == 0) return GF_OK;
  if (nseg > 1000) return gf_set_err(m->ctx, GF_ERR_ARG, "segments: at most 1000 members");
  if (eoff[0] != 0 || eoff[nseg] != (int64_t)m->nelems)
    return gf_set_err(m->ctx, GF_ERR_ARG, "segments: offsets must run from 0 to nelems");
  for (int k = 0; k < nseg; k++)
    if (eoff[k + 1] <= eoff[k]) return gf_set_err(m->ctx, GF_ERR_ARG, "segments: empty member");
  m->h_seg_eoff.assign(eoff, eoff + nseg + 1);
  // update blocks in proportion to the members' sizes, at least one each
  const int budget = std::max(nseg, 1024 - nseg);
  m->h_seg_boff.assign(nseg + 1, 0);
  for (int k = 0; k < nseg; k++) {
    const double frac = (double)(eoff[k + 1] - eoff[k]) / (double)m->nelems;
    m->h_seg_boff[k + 1] = m->h_seg_boff[k] + std::max(1, (int)(frac * budget + 0.5));
  }
  // the reduction buffer holds the apply partials (one per logical block),
  // then the update blocks' pairs, and the set-up pairs plus one pair per
  // member for the per-member reductions
  {
    const int nbU = m->h_seg_boff[nseg];
    const size_t need = std::max(std::max((size_t)col_blocks(m), m->nelems) + 64 + 2 * (size_t)nbU,
                                 2 * (size_t)nbU + 64 + 2 * (size_t)nseg) + 64;
    if (m->red_cap < need) {
      GF_HIP(m->ctx, hipFree(m->red));
      m->red = nullptr;
      GF_HIP(m->ctx, hipMalloc((void **)&m->red, need * sizeof(double)));
      m->red_cap = need;
    }
  }
  GF_HIP(m->ctx, hipMalloc((void **)&m->seg_eoff, (nseg + 1) * sizeof(int64_t)));
  GF_HIP(m->ctx, hipMalloc((void **)&m->seg_boff, (nseg + 1) * sizeof(int)));
  GF_HIP(m->ctx, hipMalloc((void **)&m->sscal, (size_t)nseg * kSegScal * sizeof(double)));
  GF_HIP(m->ctx, hipHostMalloc((void **)&m->hsscal, (size_t)nseg * kSegScal * sizeof(double),
                               hipHostMallocDefault));
  GF_HIP(m->ctx, hipMemcpy(m->seg_eoff, eoff, (nseg + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  GF_HIP(m->ctx, hipMemcpy(m->seg_boff, m->h_seg_boff.data(), (nseg + 1) * sizeof(int),
                           hipMemcpyHostToDevice));
  m->nseg = nseg;
  // no GGFX group may join two members
  double bad = 0.0;
  double *dbad = m->scal + 63;
  GF_HIP(m->ctx, hipMemcpy(dbad, &bad, sizeof(double), hipMemcpyHostToDevice));
  if (m->ngroups)
    hipLaunchKernelGGL(k_groups_in_segments, dim3(nblocks((size_t)m->ngroups)), dim3(kThreads), 0,
                       m->ctx->stream, m->goff, m->gmem, m->ngroups, m->seg_eoff, nseg, m->nn, dbad);
  GF_LAUNCH_CHECK(m->ctx);
  GF_HIP(m->ctx, hipMemcpyAsync(&bad, dbad, sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  if (bad != 0.0) {
    gf_mesh_set_segments(m, 0, nullptr);
    return gf_set_err(m->ctx, GF_ERR_ARG, "segments: a GGFX group joins two members");
  }
  return GF_OK;
}

extern "C" int gf_cg_member_results(gf_mesh *m, int nseg, int *iters, double *rnorm,
                                    double *rnorm0) {
  if (!m || nseg != m->nseg || (int)m->seg_iters.size() != nseg) return GF_ERR_ARG;
  for (int k = 0; k < nseg; k++) {
    if (iters) iters[k] = m->seg_iters[k];
    if (rnorm) rnorm[k] = m->seg_rnorm[k];
    if (rnorm0) rnorm0[k] = m->seg_rnorm0[k];
  }
  return GF_OK;
}

extern "C" int gf_cg_solve(gf_mesh *m, const gf_cg_params *prm, const double *b, const double *xb,
                           double *x, gf_cg_result *res) {
  RC(cg_check_args(m, prm, b, x, res));
  const size_t n = m->ndof, bytes = n * sizeof(double);
  CGWork wk;
  wk.r = (double *)gf_ws(m, WS_CG_BEGIN + 0, bytes);
  wk.w = (double *)gf_ws(m, WS_CG_BEGIN + 1, bytes);
  wk.q = (double *)gf_ws(m, WS_CG_BEGIN + 2, bytes);
  wk.dinv = prm->precond ? (double *)gf_ws(m, WS_CG_BEGIN + 3, bytes) : nullptr;
  wk.resid = (double *)gf_ws(m, WS_CG_BEGIN + 4, (size_t)(prm->maxit + 8) * sizeof(double));
  if (!wk.r || !wk.w || !wk.q || (prm->precond && !wk.dinv) || !wk.resid)
    return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  if (prm->precision != 0) {
    const size_t b32 = n * sizeof(float);
    wk.r32 = (float *)gf_ws(m, WS_CG_BEGIN + 5, b32);
    wk.w32 = (float *)gf_ws(m, WS_CG_BEGIN + 6, b32);
    wk.q32 = (float *)gf_ws(m, WS_CG_BEGIN + 7, b32);
    wk.x32 = (float *)gf_ws(m, WS_CG_BEGIN + 8, b32);
    wk.dinv32 = prm->precond ? (float *)gf_ws(m, WS_CG_BEGIN + 9, b32) : nullptr;
    wk.qf32 = prm->q_field ? (float *)gf_ws(m, WS_CG_BEGIN + 10, b32) : nullptr;
    if (!wk.r32 || !wk.w32 || !wk.q32 || !wk.x32 || (prm->precond && !wk.dinv32) ||
        (prm->q_field && !wk.qf32))
      return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  }
  if (m->nseg > 0) return cg_solve_seg(m, prm, b, xb, x, res, wk);
  return cg_solve_impl(m, prm, b, xb, x, res, wk);
}

// ===========================================================================
// Workspace reservation: the slots each family of entry points takes from
// the mesh workspace (gf_ws), allocated ahead of the calls.
// ===========================================================================
extern "C" int gf_mesh_reserve(gf_mesh *m, int what, const gf_reserve_params *prm) {
  if (!m) return GF_ERR_ARG;
  const size_t n = m->ndof, bytes = n * sizeof(double);
  const gf_reserve_params z = {0, 0, 0, 0, 0};
  const gf_reserve_params &p = prm ? *prm : z;
  auto want = [&](int slot, size_t b) -> int {
    if (!gf_ws(m, slot, b)) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace reservation");
    return GF_OK;
  };
  if (what & (GF_RESERVE_RK | GF_RESERVE_ADJOINT)) {
    // every RK form (fused, structured box, embedded advection, composed)
    // and the composed operators (gf_burgers_rhs, gf_advect, gf_deriv)
    for (int k = 0; k < 16; k++) RC(want(WS_RK_BEGIN + k, bytes));
    for (int k = WS_TMP0; k <= WS_TMP3; k++) RC(want(k, bytes));
  }
  if (what & (GF_RESERVE_CG | GF_RESERVE_CG_MIXED)) {
    for (int k = 0; k < 4; k++) RC(want(WS_CG_BEGIN + k, bytes));
    const int maxit = p.cg_maxit > 0 ? p.cg_maxit : 512;
    RC(want(WS_CG_BEGIN + 4, (size_t)(maxit + 8) * sizeof(double)));
  }
  if (what & GF_RESERVE_CG_MIXED)
    for (int k = 5; k <= 10; k++) RC(want(WS_CG_BEGIN + k, n * sizeof(float)));
  if (what & GF_RESERVE_ADJOINT) {
    const int nv = m->gtype == GF_GE_2DEMBEDDED ? 3 : m->dim;
    const int order = p.adj_order >= 1 && p.adj_order <= 4 ? p.adj_order : 4;
    for (int k = 0; k < 6; k++) RC(want(WS_ADJ_TMP + k, bytes));
    for (int k = 0; k < 2 * order * nv + 4 * nv; k++) RC(want(WS_ADJ_BEGIN + k, bytes));
    if (p.adj_nsteps > 0 && p.adj_ckpt > 0) {
      const int ns = p.nstate > 0 ? p.nstate : nv;
      const int ncp = (p.adj_nsteps + p.adj_ckpt - 1) / p.adj_ckpt;
      for (int k = 0; k < (ncp + p.adj_ckpt + 1) * ns; k++) RC(want(WS_ADJ_CKPT + k, bytes));
    }
  }
  if (what & GF_RESERVE_FREEZE) m->ws_frozen = true;
  GF_HIP(m->ctx, hipDeviceSynchronize());
  return GF_OK;
}

// ===========================================================================
// Diagnostics: time one kernel of the path with HIP events on the context
// stream.  The kernel is launched `reps` times back to back on scratch data
// of the mesh's size; the average duration per launch is returned in ms.
// ===========================================================================
namespace {
__global__ void k_fill_pattern(double *a, size_t n, double s) {
  GRID_LOOP(g, n) a[g] = s * (double)((g * 2654435761u) % 1000u) * 1e-3;
}
__global__ void k_fill_pattern_f(float *a, size_t n, float s) {
  GRID_LOOP(g, n) a[g] = s * (float)((g * 2654435761u) % 1000u) * 1e-3f;
}
}  // namespace

extern "C" int gf_time_kernel(gf_mesh *m, const char *which, int reps, double *avg_ms) {
  if (!m || !which || reps < 1 || !avg_ms) return GF_ERR_ARG;
  if (m->gtype != GF_GE_REGULAR || !m->Niso)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "timing: regular isotropic meshes");
  std::string w(which);
  hipStream_t s = m->ctx->stream;
  size_t n = m->ndof, bytes = n * sizeof(double);
  const int nbuf = 16;
  double *buf[nbuf];
  for (int k = 0; k < nbuf; k++) {
    buf[k] = (double *)gf_ws(m, WS_SLOTS + k, bytes);
    if (!buf[k]) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
    hipLaunchKernelGGL(k_fill_pattern, dim3(nblocks(n)), dim3(kThreads), 0, s, buf[k], n,
                       0.5 + 0.1 * k);
  }
  GF_LAUNCH_CHECK(m->ctx);
  std::function<int()> launch;
  StageArgs sa;
  CGApplyArgs ca;
  memset(&sa, 0, sizeof(sa));
  memset(&ca, 0, sizeof(ca));
  CGApply3Args ca3;
  ca3.seg_eoff = nullptr;
  ca3.nseg = 0;
  ca3.sscal = nullptr;
  memset(&ca3, 0, sizeof(ca3));
  const bool pencil = w == "burgers_stage" || w == "cg_apply" || w == "cg_apply_f32" ||
                      w == "burgers_stage_box";
  if (pencil && m->dim == 3 && !m->dconst)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "pencil kernels need the standard GLL D");
  if (w == "burgers_stage") {
    // stages 2-3 form: reads uin, kprev, accsrc; writes acc, kout (5 fields per component)
    for (int c = 0; c < m->dim; c++) {
      sa.uin[c] = buf[c];
      sa.kprev[c] = buf[3 + c];
      sa.acc[c] = buf[6 + c];
      sa.accsrc[c] = buf[6 + c];
      sa.kout[c] = buf[9 + c];
    }
    sa.bh = 0.5e-5;
    sa.ch = 1e-5 / 3.0;
    sa.D = m->D[0];
    sa.W = m->W[0];
    sa.edx = m->edx;
    sa.ejac = m->ejac;
    sa.imass = m->imass;
    sa.nu = 0.02;
    sa.nelems = m->nelems;
    if (m->dim == 3)
      launch = [&]() { return GF_DISPATCH_N3(m->Niso, launch_stage3p, m, sa); };
    else
      launch = [&]() { return GF_DISPATCH_N(m->dim, m->Niso, launch_stage, m, sa); };
  } else if (w == "cg_apply" || w == "cg_apply_f32") {
    if (m->dim != 3) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "cg_apply timing: 3-D meshes");
    bool f32 = w == "cg_apply_f32";
    double hs[12] = {0};
    hs[S_BETA] = 0.5;
    hs[S_FIRST] = 0.0;
    hs[S_DONE] = 0.0;
    GF_HIP(m->ctx, hipMemcpyAsync(m->scal, hs, sizeof(hs), hipMemcpyHostToDevice, s));
    if (f32) {
      for (int k = 0; k < 4; k++)
        hipLaunchKernelGGL(k_fill_pattern_f, dim3(nblocks(n)), dim3(kThreads), 0, s,
                           (float *)buf[k], n, 0.5f + 0.1f * k);
    }
    ca3.r = buf[0];
    ca3.dinv = buf[1];
    ca3.dtab = nullptr;
    ca3.deq = nullptr;
    ca3.w = buf[2];
    ca3.q = buf[3];
    ca3.D = f32 ? (const void *)m->D32 : (const void *)m->D[0];
    for (int c = 0; c < 3; c++) {
      ca3.DR[c] = f32 ? (const void *)m->pDR32[c] : (const void *)m->pDR[c];
      ca3.DC[c] = f32 ? (const void *)m->pDC32[c] : (const void *)m->pDC[c];
    }
    ca3.W = f32 ? (const void *)m->W32 : (const void *)m->W[0];
    ca3.edx = f32 ? (const void *)m->edx32 : (const void *)m->edx;
    ca3.ejac = f32 ? (const void *)m->ejac32 : (const void *)m->ejac;
    ca3.qf = nullptr;
    ca3.qs = 1.0;
    ca3.mask = nullptr;
    ca3.topo = m->box;
    ca3.tmask = (m->has_mask && m->has_box_mask) ? 1 : 0;
    ca3.p = 0.1;
    ca3.scal = m->scal;
    ca3.part = m->red;
    ca3.nelems = m->nelems;
    if (f32)
      launch = [&]() { return GF_DISPATCH_N3(m->Niso, cgapply3c_f, m, ca3); };
    else
      launch = [&]() { return GF_DISPATCH_N3(m->Niso, cgapply3c_d, m, ca3); };
  } else if (w == "rk_out_box") {
    // structured-box step output: reads acc and K (both gathered), writes u
    if (m->dim != 3 || !m->has_box)
      return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "rk_out_box timing: structured 3-D boxes");
    launch = [&]() {
      return GF_DISPATCH_N3(m->Niso, launch_rk_out_box, m, buf, buf + 6, buf + 9, 1e-5 / 6.0);
    };
  } else if (w == "burgers_stage_box") {
    // structured-box stage, middle-stage form: reads uin, K_prev (gathered),
    // acc; writes acc, K (5 fields per component)
    if (m->dim != 3 || !m->has_box)
      return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "burgers_stage_box timing: structured 3-D boxes");
    launch = [&]() {
      return GF_DISPATCH_N3(m->Niso, launch_stage3b, m, buf, buf + 3, buf + 6, buf + 9, 0.5e-5,
                            1e-5 / 3.0, 0.02, 2, 0);
    };
  } else if (w == "cg_update" || w == "cg_update_f32") {
    // CG update with the Jacobi preconditioner: x, w, r, q (the diagonal
    // formed in the kernel on a uniform box, else read from dinv)
    if (m->dim != 3 || !m->has_box)
      return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "cg_update timing: structured 3-D boxes");
    bool f32 = w == "cg_update_f32";
    double hs[S_COUNT] = {0};
    hs[S_ALPHA] = 1e-3;
    GF_HIP(m->ctx, hipMemcpyAsync(m->scal, hs, sizeof(hs), hipMemcpyHostToDevice, s));
    if (f32) {
      for (int k = 0; k < 5; k++)
        hipLaunchKernelGGL(k_fill_pattern_f, dim3(nblocks(n)), dim3(kThreads), 0, s,
                           (float *)buf[k], n, 0.5f + 0.1f * k);
      launch = [&]() {
        CGVecs<float> v{};
        v.x = (float *)buf[0];
        v.w = (float *)buf[1];
        v.r = (float *)buf[2];
        v.q = (float *)buf[3];
        v.dinv = (const float *)buf[4];
        v.W = m->W32;
        v.edx = m->edx32;
        v.ejac = m->ejac32;
        v.p = 0.1;
        v.qs = 1.0;
        return launch_update_box<float, GF_NORM_ASSEMBLED_L2>(m, v, m->red);
      };
    } else {
      launch = [&]() {
        CGVecs<double> v{};
        v.x = buf[0];
        v.w = buf[1];
        v.r = buf[2];
        v.q = buf[3];
        v.dinv = buf[4];
        v.W = m->W[0];
        v.edx = m->edx;
        v.ejac = m->ejac;
        v.p = 0.1;
        v.qs = 1.0;
        return launch_update_box<double, GF_NORM_ASSEMBLED_L2>(m, v, m->red);
      };
    }
  } else if (w == "helmholtz") {
    launch = [&]() { return gf_helm_t(m, 0.1, nullptr, nullptr, 1.0, buf[0], buf[1]); };
  } else {
    return gf_set_err(m->ctx, GF_ERR_ARG, "unknown kernel name");
  }
  RC(launch());  // warm-up
  hipEvent_t e0, e1;
  GF_HIP(m->ctx, hipEventCreate(&e0));
  GF_HIP(m->ctx, hipEventCreate(&e1));
  GF_HIP(m->ctx, hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) RC(launch());
  GF_HIP(m->ctx, hipEventRecord(e1, s));
  GF_HIP(m->ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  GF_HIP(m->ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / reps;
  return GF_OK;
}
