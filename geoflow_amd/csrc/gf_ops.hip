// gf_ops.hip -- context, mesh and the element-local operators of the SEM path:
// tensor-product derivatives, gradient, divergence, mass, Helmholtz, DSS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "gf_dtab.hpp"
#include "gf_kernels.hpp"

using gf::kThreads;
using gf::Shape;

// Scratch of gf_integrate_noalloc: the per-block partial sums of the
// integral, one slot per block of its partial-sum kernel.  Each device the
// library runs on holds its own copy until the process ends; host code uses
// the per-device address that gf_ctx_create takes with hipGetSymbolAddress
// (gf_ctx::scratch), never this name.
constexpr int kNoAllocParts = 1024;
static __device__ __align__(256) double scratch[kNoAllocParts];

// ===========================================================================
// Context, errors, workspace
// ===========================================================================
int gf_set_err(gf_ctx *ctx, int code, const char *msg) {
  if (ctx) ctx->err = msg ? msg : "";
  return code;
}

int gf_check_hip(gf_ctx *ctx, hipError_t e, const char *what) {
  if (ctx) {
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
  }
  return e == hipErrorOutOfMemory ? GF_ERR_ALLOC : GF_ERR_HIP;
}

size_t gf_red_need(size_t nelems) {
  // apply partials (one per block, at most one block per element) at the
  // front, then the update kernels' (rho, norm) pairs and the set-up slots
  const size_t need = nelems + 64 + 2 * 1024 + 128;
  const size_t base = (size_t)GF_RED_MAX_BLOCKS * 8;
  return need > base ? need : base;
}

void *gf_ws(gf_mesh *m, int slot, size_t bytes) {
  if ((int)m->ws.size() <= slot) {
    if (m->ws_frozen) return nullptr;
    m->ws.resize(slot + 1, nullptr);
    m->ws_cap.resize(slot + 1, 0);
  }
  if (m->ws[slot] && m->ws_cap[slot] >= bytes) return m->ws[slot];
  if (m->ws_frozen) return nullptr;  // reserved workspace only
  if (m->ws[slot]) (void)hipFree(m->ws[slot]);
  m->ws[slot] = nullptr;
  m->ws_cap[slot] = 0;
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  m->ws[slot] = p;
  m->ws_cap[slot] = bytes;
  return p;
}

extern "C" {

const char *gf_version(void) { return "geoflow_amd 0.1 (gfx950)"; }

int gf_ctx_create(int device, void *stream, gf_ctx **out) {
  if (!out) return GF_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return GF_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return GF_ERR_HIP;
  gf_ctx *c = new gf_ctx_s();
  c->device = device;
  c->stream = (hipStream_t)stream;
  // this device's copy of the scratch array (with the device current)
  void *sp = nullptr;
  hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !sp) {
    delete c;
    return GF_ERR_HIP;
  }
  c->scratch = (double *)sp;
  *out = c;
  return GF_OK;
}

int gf_ctx_destroy(gf_ctx *ctx) {
  delete ctx;
  return GF_OK;
}

int gf_ctx_set_stream(gf_ctx *ctx, void *stream) {
  if (!ctx) return GF_ERR_ARG;
  ctx->stream = (hipStream_t)stream;
  return GF_OK;
}

int gf_ctx_synchronize(gf_ctx *ctx) {
  if (!ctx) return GF_ERR_ARG;
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

const char *gf_ctx_last_error(gf_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int gf_malloc(gf_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return GF_ERR_ARG;
  GF_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
  return GF_OK;
}
int gf_free(gf_ctx *ctx, void *dptr) {
  if (!ctx) return GF_ERR_ARG;
  GF_HIP(ctx, hipFree(dptr));
  return GF_OK;
}
int gf_memcpy_h2d(gf_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return GF_ERR_ARG;
  GF_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}
int gf_memcpy_d2h(gf_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return GF_ERR_ARG;
  GF_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

}  // extern "C"

// ===========================================================================
// Kernels
// ===========================================================================
namespace {
// Entry (l,m) of the compile-time GLL derivative matrix for N points.
double gll_const(int N, int l, int m) {
  switch (N) {
    case 2: return gf::GllD<2>::at(l, m);
    case 3: return gf::GllD<3>::at(l, m);
    case 4: return gf::GllD<4>::at(l, m);
    case 5: return gf::GllD<5>::at(l, m);
    case 6: return gf::GllD<6>::at(l, m);
    case 7: return gf::GllD<7>::at(l, m);
    case 8: return gf::GllD<8>::at(l, m);
    case 9: return gf::GllD<9>::at(l, m);
    case 10: return gf::GllD<10>::at(l, m);
    default: return 0.0;
  }
}
}  // namespace

namespace {

// ---- derivative along one direction (isotropic N) -------------------------
// du = Dop_dir u, optionally times the per-element physical factor
// edx[e*DIM+dir] applied after the sum (GGrid::deriv, ggrid.ipp:1175-1177).
template <int DIM, int N, bool TRANS>
__global__ __launch_bounds__(kThreads) void k_deriv(const double *__restrict__ u,
                                                    double *__restrict__ du,
                                                    const double *__restrict__ D,
                                                    const double *__restrict__ edx, int dir,
                                                    size_t nelems) {
  using S = Shape<DIM, N>;
  __shared__ double su[S::NODES];
  __shared__ double sD[N * N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  gf::stage(su, u, base, nel * S::NN);
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(r, i, j, k);
    double v = gf::contract<DIM, N, TRANS>(su + el * S::NN, sD, dir, i, j, k);
    if (edx) v = v * edx[(e0 + el) * DIM + dir];
    du[base + q] = v;
  }
}

// ---- all directions at once (compute_grefderivs / grad) -------------------
template <int DIM, int N>
__global__ __launch_bounds__(kThreads) void k_grad(const double *__restrict__ u, double *d0,
                                                   double *d1, double *d2,
                                                   const double *__restrict__ D,
                                                   const double *__restrict__ edx, size_t nelems) {
  using S = Shape<DIM, N>;
  __shared__ double su[S::NODES];
  __shared__ double sD[N * N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  gf::stage(su, u, base, nel * S::NN);
  __syncthreads();
  double *outs[3] = {d0, d1, d2};
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(r, i, j, k);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double v = gf::contract<DIM, N, false>(su + el * S::NN, sD, d, i, j, k);
      if (edx) v = v * edx[(e0 + el) * DIM + d];
      outs[d][base + q] = v;
    }
  }
}

// ---- reference-space divergence (compute_grefdiv) --------------------------
// divu = ((0 + D1 u1) + D2 u2) + D3 u3 (ggrid.ipp:1870-1925); NULL inputs skipped.
template <int DIM, int N, bool TRANS>
__global__ __launch_bounds__(kThreads) void k_div(const double *__restrict__ u0,
                                                  const double *__restrict__ u1,
                                                  const double *__restrict__ u2,
                                                  double *__restrict__ out,
                                                  const double *__restrict__ D, size_t nelems) {
  using S = Shape<DIM, N>;
  __shared__ double su[S::NODES];
  __shared__ double sD[N * N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  const double *ins[3] = {u0, u1, u2};
  double acc[(S::NODES + kThreads - 1) / kThreads];
#pragma unroll
  for (int t = 0; t < (S::NODES + kThreads - 1) / kThreads; t++) acc[t] = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    if (!ins[d]) continue;
    __syncthreads();
    gf::stage(su, ins[d], base, nel * S::NN);
    __syncthreads();
    int t = 0;
    for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x, t++) {
      int el = q / S::NN, r = q - el * S::NN, i, j, k;
      gf::node_ijk<DIM, N>(r, i, j, k);
      acc[t] += gf::contract<DIM, N, TRANS>(su + el * S::NN, sD, d, i, j, k);
    }
  }
  int t = 0;
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x, t++) out[base + q] = acc[t];
}

// ---- regular-element Helmholtz ------------------------------------------
// uo = p * sum_k D_k^T ( M * (G_kk * (D_k u)) [* p_field] )  +  q * M * u
// with G_kk = (dXi_k/dX_k)^2 (reg_init, ghelmholtz.ipp:571-575) and the lumped
// mass computed in-kernel from the weights and the element Jacobian
// (reg_prod, ghelmholtz.ipp:308-378).
template <int DIM, int N>
__global__ __launch_bounds__(kThreads) void k_helm(const double *__restrict__ u,
                                                   double *__restrict__ uo,
                                                   const double *__restrict__ D,
                                                   const double *__restrict__ W,
                                                   const double *__restrict__ edx,
                                                   const double *__restrict__ ejac, double pscal,
                                                   const double *__restrict__ pfield,
                                                   const double *__restrict__ qfield,
                                                   double qscal, size_t nelems) {
  using S = Shape<DIM, N>;
  __shared__ double su[S::NODES];
  __shared__ double sg[DIM][S::NODES];
  __shared__ double sD[N * N];
  __shared__ double sW[N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  if (threadIdx.x < N) sW[threadIdx.x] = W[threadIdx.x];
  gf::stage(su, u, base, nel * S::NN);
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(r, i, j, k);
    size_t e = e0 + el;
    double M = gf::node_mass<DIM>(sW, i, j, k, ejac[e]);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      double g = gf::contract<DIM, N, false>(su + el * S::NN, sD, d, i, j, k);
      double f = edx[e * DIM + d];
      g = g * (f * f);
      if (pfield) g = g * pfield[base + q];
      sg[d][q] = M * g;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<DIM, N>(r, i, j, k);
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) acc += gf::contract<DIM, N, true>(&sg[d][el * S::NN], sD, d, i, j, k);
    if (!pfield && pscal != 1.0) acc *= pscal;
    if (qfield || qscal != 0.0) {
      double M = gf::node_mass<DIM>(sW, i, j, k, ejac[e0 + el]);
      double t = M * su[q];
      if (qfield) t *= qfield[base + q];
      else if (qscal != 1.0) t *= qscal;
      acc += t;
    }
    uo[base + q] = acc;
  }
}

// ---- 2-D embedded Helmholtz (embed_prod, ghelmholtz.ipp:205-290) -----------
// gdu_j = sum_i G(i,j) D_i u ; uo = p sum_j D_j^T gdu_j + q M u
template <int DIM, int N>
__global__ __launch_bounds__(kThreads) void k_helm_embed(
    const double *__restrict__ u, double *__restrict__ uo, const double *__restrict__ D,
    const double *__restrict__ G00, const double *__restrict__ G01, const double *__restrict__ G11,
    const double *__restrict__ mass, double pscal, const double *__restrict__ pfield,
    const double *__restrict__ qfield, double qscal, size_t nelems) {
  using S = Shape<2, N>;
  __shared__ double su[S::NODES];
  __shared__ double sg[2][S::NODES];
  __shared__ double sD[N * N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  gf::stage(su, u, base, nel * S::NN);
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<2, N>(r, i, j, k);
    double d0 = gf::contract<2, N, false>(su + el * S::NN, sD, 0, i, j, k);
    double d1 = gf::contract<2, N, false>(su + el * S::NN, sD, 1, i, j, k);
    size_t g = base + q;
    double g0 = 0.0 + d0 * G00[g];
    g0 += d1 * G01[g];
    double g1 = 0.0 + d0 * G01[g];
    g1 += d1 * G11[g];
    if (pfield) {
      g0 *= pfield[g];
      g1 *= pfield[g];
    }
    sg[0][q] = g0;
    sg[1][q] = g1;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<2, N>(r, i, j, k);
    double acc = 0.0;
    acc += gf::contract<2, N, true>(&sg[0][el * S::NN], sD, 0, i, j, k);
    acc += gf::contract<2, N, true>(&sg[1][el * S::NN], sD, 1, i, j, k);
    if (!pfield && pscal != 1.0) acc *= pscal;
    if (qfield || qscal != 0.0) {
      double t = mass[base + q] * su[q];
      if (qfield) t *= qfield[base + q];
      else if (qscal != 1.0) t *= qscal;
      acc += t;
    }
    uo[base + q] = acc;
  }
}

// ---- 3-D deformed Helmholtz (def_prod, ghelmholtz.ipp:108-191) ------------
// gdu_j = ((0 + G(0,j) D1 u) + G(1,j) D2 u) + G(2,j) D3 u ;
// uo = p (((0 + D1^T gdu_0) + D2^T gdu_1) + D3^T gdu_2) + q M u,
// with the six stored metric vectors (G00, G01, G11, G02, G12, G22).
template <int N>
__global__ __launch_bounds__(kThreads) void k_helm_def3(
    const double *__restrict__ u, double *__restrict__ uo, const double *__restrict__ D,
    const double *__restrict__ G00, const double *__restrict__ G01,
    const double *__restrict__ G11, const double *__restrict__ G02,
    const double *__restrict__ G12, const double *__restrict__ G22,
    const double *__restrict__ mass, double pscal, const double *__restrict__ pfield,
    const double *__restrict__ qfield, double qscal, size_t nelems) {
  using S = Shape<3, N>;
  __shared__ double su[S::NODES];
  __shared__ double sg[3][S::NODES];
  __shared__ double sD[N * N];
  size_t e0 = (size_t)blockIdx.x * S::EPB;
  int nel = (int)min((size_t)S::EPB, nelems - e0);
  size_t base = e0 * S::NN;
  gf::stage_D(sD, D, N);
  gf::stage(su, u, base, nel * S::NN);
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<3, N>(r, i, j, k);
    const double d0 = gf::contract<3, N, false>(su + el * S::NN, sD, 0, i, j, k);
    const double d1 = gf::contract<3, N, false>(su + el * S::NN, sD, 1, i, j, k);
    const double d2 = gf::contract<3, N, false>(su + el * S::NN, sD, 2, i, j, k);
    const size_t g = base + q;
    const double a00 = G00[g], a01 = G01[g], a11 = G11[g], a02 = G02[g], a12 = G12[g],
                 a22 = G22[g];
    double g0 = 0.0 + d0 * a00;
    g0 += d1 * a01;
    g0 += d2 * a02;
    double g1 = 0.0 + d0 * a01;
    g1 += d1 * a11;
    g1 += d2 * a12;
    double g2 = 0.0 + d0 * a02;
    g2 += d1 * a12;
    g2 += d2 * a22;
    if (pfield) {
      g0 *= pfield[g];
      g1 *= pfield[g];
      g2 *= pfield[g];
    }
    sg[0][q] = g0;
    sg[1][q] = g1;
    sg[2][q] = g2;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nel * S::NN; q += blockDim.x) {
    int el = q / S::NN, r = q - el * S::NN, i, j, k;
    gf::node_ijk<3, N>(r, i, j, k);
    double acc = 0.0;
    acc += gf::contract<3, N, true>(&sg[0][el * S::NN], sD, 0, i, j, k);
    acc += gf::contract<3, N, true>(&sg[1][el * S::NN], sD, 1, i, j, k);
    acc += gf::contract<3, N, true>(&sg[2][el * S::NN], sD, 2, i, j, k);
    if (!pfield && pscal != 1.0) acc *= pscal;
    if (qfield || qscal != 0.0) {
      double t = mass[base + q] * su[q];
      if (qfield) t *= qfield[base + q];
      else if (qscal != 1.0) t *= qscal;
      acc += t;
    }
    uo[base + q] = acc;
  }
}

// ---- generic (anisotropic) derivative: one thread per node, operands from
// global memory through the cache.
__global__ void k_deriv_generic(const double *__restrict__ u, double *__restrict__ du,
                                const double *__restrict__ D, int dir, int trans, int N0, int N1,
                                int N2, size_t ndof, const double *__restrict__ edx, int dim) {
  size_t nn = (size_t)N0 * N1 * N2;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < ndof;
       g += (size_t)gridDim.x * blockDim.x) {
    size_t e = g / nn;
    int r = (int)(g - e * nn);
    int i = r % N0, j = (r / N0) % N1, k = r / (N0 * N1);
    const double *ue = u + e * nn;
    double acc = 0.0;
    if (dir == 0) {
      for (int m = 0; m < N0; m++) {
        double d = trans ? D[m + i * N0] : D[i + m * N0];
        acc = fma(d, ue[(size_t)k * N0 * N1 + j * N0 + m], acc);
      }
    } else if (dir == 1) {
      for (int m = 0; m < N1; m++) {
        double d = trans ? D[m + j * N1] : D[j + m * N1];
        acc = fma(d, ue[(size_t)k * N0 * N1 + m * N0 + i], acc);
      }
    } else {
      for (int m = 0; m < N2; m++) {
        double d = trans ? D[m + k * N2] : D[k + m * N2];
        acc = fma(d, ue[(size_t)m * N0 * N1 + j * N0 + i], acc);
      }
    }
    if (edx) acc = acc * edx[e * dim + dir];
    du[g] = acc;
  }
}

// ---- mass -----------------------------------------------------------------
template <int DIM, int N>
__global__ void k_mass_init(double *__restrict__ mass, const double *__restrict__ W,
                            const double *__restrict__ ejac, size_t ndof) {
  constexpr int NN = Shape<DIM, N>::NN;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < ndof;
       g += (size_t)gridDim.x * blockDim.x) {
    size_t e = g / NN;
    int r = (int)(g - e * NN), i, j, k;
    gf::node_ijk<DIM, N>(r, i, j, k);
    mass[g] = gf::node_mass<DIM>(W, i, j, k, ejac[e]);
  }
}

__global__ void k_mass_init_generic(double *__restrict__ mass, const double *__restrict__ W0,
                                    const double *__restrict__ W1, const double *__restrict__ W2,
                                    const double *__restrict__ jac, int pernode, int N0, int N1,
                                    int N2, int dim, size_t ndof) {
  size_t nn = (size_t)N0 * N1 * N2;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < ndof;
       g += (size_t)gridDim.x * blockDim.x) {
    size_t e = g / nn;
    int r = (int)(g - e * nn);
    int i = r % N0, j = (r / N0) % N1, k = r / (N0 * N1);
    double w = dim == 3 ? W2[k] * W1[j] * W0[i] : W1[j] * W0[i];
    mass[g] = w * (pernode ? jac[g] : jac[e]);
  }
}

__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b,
                      double *__restrict__ out, size_t n) {
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n;
       g += (size_t)gridDim.x * blockDim.x)
    out[g] = a[g] * b[g];
}

__global__ void k_recip(double *__restrict__ a, size_t n) {
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n;
       g += (size_t)gridDim.x * blockDim.x)
    a[g] = 1.0 / a[g];
}

__global__ void k_narrow_f32(float *__restrict__ o, const double *__restrict__ a, size_t n) {
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n;
       g += (size_t)gridDim.x * blockDim.x)
    o[g] = (float)a[g];
}

// ---- DSS over coincidence groups --------------------------------------------
// Each thread owns one group with more than one member; members are visited
// in ascending local id and accumulated from zero, the reference's
// std::accumulate order (ggfx.hpp:60-76, 401-405), then written back to every
// member.  Singleton groups are identities and are skipped.
// The grid is a multiple of 8 and its blocks are renumbered so that each of
// the 8 XCDs (each with its own L2) walks one contiguous range of groups:
// groups of neighbouring indices share elements, and so cache lines.
template <int OP, typename T>
__global__ void k_dss(T *__restrict__ u, const int *__restrict__ goff,
                      const int *__restrict__ gmem, int ngroups) {
  const int bx = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  for (int g = bx * blockDim.x + threadIdx.x; g < ngroups; g += gridDim.x * blockDim.x) {
    int b = goff[g], e = goff[g + 1];
    if (e - b <= 8) {
      // groups of at most 8 (boxes, the sphere): every member index and
      // value is loaded before the first use, so a thread waits for memory
      // twice instead of once per member; same order, same result
      const int cnt = e - b;
      int id[8];
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) id[k] = k < cnt ? gmem[b + k] : 0;
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = k < cnt ? u[id[k]] : T(0);
      T r;
      if (OP == GF_DSS_SUM || OP == GF_DSS_SMOOTH) {
        r = T(0);
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (k < cnt) r = r + v[k];
        if (OP == GF_DSS_SMOOTH) r = r / (T)cnt;
      } else if (OP == GF_DSS_MAX) {
        r = std::numeric_limits<T>::lowest();
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (k < cnt) r = fmax(r, v[k]);
      } else {
        r = std::numeric_limits<T>::max();
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (k < cnt) r = fmin(r, v[k]);
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < cnt) u[id[k]] = r;
      continue;
    }
    T r;
    if (OP == GF_DSS_SUM || OP == GF_DSS_SMOOTH) {
      r = T(0);
      for (int s = b; s < e; s++) r = r + u[gmem[s]];
      if (OP == GF_DSS_SMOOTH) r = r / (T)(e - b);
    } else if (OP == GF_DSS_MAX) {
      r = std::numeric_limits<T>::lowest();
      for (int s = b; s < e; s++) r = fmax(r, u[gmem[s]]);
    } else {
      r = std::numeric_limits<T>::max();
      for (int s = b; s < e; s++) r = fmin(r, u[gmem[s]]);
    }
    for (int s = b; s < e; s++) u[gmem[s]] = r;
  }
}

// mult/imult from groups: start from 1, overwrite group members.
__global__ void k_fill(double *__restrict__ a, double v, size_t n) {
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n;
       g += (size_t)gridDim.x * blockDim.x)
    a[g] = v;
}
__global__ void k_imult_groups(double *__restrict__ imult, const int *__restrict__ goff,
                               const int *__restrict__ gmem, int ngroups) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += gridDim.x * blockDim.x) {
    int b = goff[g], e = goff[g + 1];
    double v = 1.0 / (double)(e - b);
    for (int s = b; s < e; s++) imult[gmem[s]] = v;
  }
}

// Deterministic two-stage sum: per-block partials then one block in order.
__global__ void k_partial_dot(const double *__restrict__ a, const double *__restrict__ b,
                              double *__restrict__ part, size_t n) {
  __shared__ double sred[kThreads / 64];
  double v = 0.0;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n;
       g += (size_t)gridDim.x * blockDim.x)
    v += a[g] * (b ? b[g] : 1.0);
  double r = gf::block_reduce<false>(v, sred);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}
__global__ void k_final_sum(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double sred[kThreads / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
  double r = gf::block_reduce<false>(v, sred);
  if (threadIdx.x == 0) *out = r;
}

inline int nblocks(size_t n, int per = kThreads) {
  size_t b = (n + per - 1) / per;
  if (b > 65536) b = 65536;
  return (int)std::max<size_t>(b, 1);
}

}  // namespace

// ===========================================================================
// Launch wrappers (templated on DIM, N)
// ===========================================================================
namespace gfl {

template <int DIM, int N>
int launch_deriv(gf_mesh *m, int dir, int trans, const double *u, double *du, const double *edx) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  if (trans)
    hipLaunchKernelGGL((k_deriv<DIM, N, true>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, du,
                       m->D[dir], edx, dir, m->nelems);
  else
    hipLaunchKernelGGL((k_deriv<DIM, N, false>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u,
                       du, m->D[dir], edx, dir, m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int DIM, int N>
int launch_grad(gf_mesh *m, const double *u, double *d0, double *d1, double *d2,
                const double *edx) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  hipLaunchKernelGGL((k_grad<DIM, N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, d0, d1, d2,
                     m->D[0], edx, m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int DIM, int N>
int launch_div(gf_mesh *m, int trans, const double *u0, const double *u1, const double *u2,
               double *out) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  if (trans)
    hipLaunchKernelGGL((k_div<DIM, N, true>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u0, u1,
                       u2, out, m->D[0], m->nelems);
  else
    hipLaunchKernelGGL((k_div<DIM, N, false>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u0, u1,
                       u2, out, m->D[0], m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int DIM, int N>
int launch_helm(gf_mesh *m, double p, const double *pf, const double *qf, double qs,
                const double *u, double *uo) {
  using S = Shape<DIM, N>;
  int nb = (int)((m->nelems + S::EPB - 1) / S::EPB);
  if (m->gtype == GF_GE_REGULAR) {
    hipLaunchKernelGGL((k_helm<DIM, N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, uo,
                       m->D[0], m->W[0], m->edx, m->ejac, p, pf, qf, qs, m->nelems);
  } else if (DIM == 2) {
    hipLaunchKernelGGL((k_helm_embed<2, N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, uo,
                       m->D[0], m->G[0], m->G[1], m->G[2], m->mass, p, pf, qf, qs, m->nelems);
  } else {
    hipLaunchKernelGGL((k_helm_def3<N>), dim3(nb), dim3(kThreads), 0, m->ctx->stream, u, uo,
                       m->D[0], m->G[0], m->G[1], m->G[2], m->G[3], m->G[4], m->G[5], m->mass, p,
                       pf, qf, qs, m->nelems);
  }
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int DIM, int N>
int launch_mass_init(gf_mesh *m) {
  hipLaunchKernelGGL((k_mass_init<DIM, N>), dim3(nblocks(m->ndof)), dim3(kThreads), 0,
                     m->ctx->stream, m->mass, m->W[0], m->ejac, m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

}  // namespace gfl

// Internal entry points used by other translation units.
int gf_dss_t(gf_mesh *m, int op, double *u) {
  if (m->ngroups == 0) return GF_OK;
  const int nb = (nblocks((size_t)m->ngroups) + 7) / 8 * 8;  // k_dss: a multiple of 8
  hipStream_t s = m->ctx->stream;
  switch (op) {
    case GF_DSS_SUM: hipLaunchKernelGGL((k_dss<GF_DSS_SUM, double>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups); break;
    case GF_DSS_SMOOTH: hipLaunchKernelGGL((k_dss<GF_DSS_SMOOTH, double>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups); break;
    case GF_DSS_MAX: hipLaunchKernelGGL((k_dss<GF_DSS_MAX, double>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups); break;
    case GF_DSS_MIN: hipLaunchKernelGGL((k_dss<GF_DSS_MIN, double>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups); break;
    default: return GF_ERR_ARG;
  }
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int gf_dss_f32(gf_mesh *m, int op, float *u) {
  if (m->ngroups == 0) return GF_OK;
  const int nb = (nblocks((size_t)m->ngroups) + 7) / 8 * 8;  // k_dss: a multiple of 8
  hipStream_t s = m->ctx->stream;
  if (op == GF_DSS_SUM)
    hipLaunchKernelGGL((k_dss<GF_DSS_SUM, float>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups);
  else if (op == GF_DSS_SMOOTH)
    hipLaunchKernelGGL((k_dss<GF_DSS_SMOOTH, float>), dim3(nb), dim3(kThreads), 0, s, u, m->goff, m->gmem, m->ngroups);
  else
    return GF_ERR_ARG;
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int gf_helm_t(gf_mesh *m, double p, const double *pf, const double *qf, double qs,
              const double *u, double *uo) {
  if (!m->Niso) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "Helmholtz needs isotropic order");
  if (qf) qs = 0.0;
  if (m->dim == 3 && m->gtype == GF_GE_REGULAR && m->dconst) return gf_helm3c(m, p, pf, qf, qs, u, uo);
  return GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_helm, m, p, pf, qf, qs, u, uo);
}

int gf_grefderiv_t(gf_mesh *m, int idir, int trans, const double *u, double *du,
                   const double *edx) {
  int dir = idir - 1;
  if (m->Niso) return GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_deriv, m, dir, trans, u, du, edx);
  hipLaunchKernelGGL(k_deriv_generic, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream,
                     u, du, m->D[dir], dir, trans, m->N[0], m->N[1], m->dim == 3 ? m->N[2] : 1,
                     m->ndof, edx, m->dim);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int gf_mul_t(gf_mesh *m, const double *a, const double *b, double *out) {
  hipLaunchKernelGGL(k_mul, dim3(nblocks(m->ndof)), dim3(kThreads), 0, m->ctx->stream, a, b, out,
                     m->ndof);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// Deterministic sum of a[i]*b[i] (b may be NULL); synchronous.
int gf_dot_t(gf_mesh *m, const double *a, const double *b, size_t n, double *result) {
  int nb = std::min(nblocks(n), 1024);
  double *part = m->red;
  hipLaunchKernelGGL(k_partial_dot, dim3(nb), dim3(kThreads), 0, m->ctx->stream, a, b, part, n);
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kThreads), 0, m->ctx->stream, part, nb,
                     m->scal + 63);
  GF_LAUNCH_CHECK(m->ctx);
  GF_HIP(m->ctx, hipMemcpyAsync(result, m->scal + 63, sizeof(double), hipMemcpyDeviceToHost,
                                m->ctx->stream));
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return GF_OK;
}

// ===========================================================================
// Mesh
// ===========================================================================
namespace {
template <typename T>
int upload(gf_ctx *c, T **d, const T *h, size_t n) {
  *d = nullptr;
  if (!h) return GF_OK;
  GF_HIP(c, hipMalloc((void **)d, n * sizeof(T) + 16));
  GF_HIP(c, hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return GF_OK;
}
}  // namespace

extern "C" {

int gf_mesh_create(gf_ctx *ctx, const gf_mesh_desc *d, gf_mesh **out) {
  if (!ctx || !d || !out) return GF_ERR_ARG;
  if (d->dim != 2 && d->dim != 3) return gf_set_err(ctx, GF_ERR_ARG, "dim must be 2 or 3");
  if (d->gtype != GF_GE_REGULAR && d->gtype != GF_GE_2DEMBEDDED && d->gtype != GF_GE_DEFORMED)
    return gf_set_err(ctx, GF_ERR_ARG, "unknown element type");
  if (d->gtype == GF_GE_2DEMBEDDED && d->dim != 2)
    return gf_set_err(ctx, GF_ERR_ARG, "embedded meshes are 2-D");
  // device-built maps: a 3-D regular isotropic box with its structure given
  const bool devmaps = d->box_device_maps != 0;
  const bool boxunion = devmaps && d->box_count >= 1 && d->box_list;
  if (devmaps && !(d->dim == 3 && d->gtype == GF_GE_REGULAR &&
                   (boxunion ? d->box_list != nullptr : d->box_ne[0] > 0) &&
                   d->N[0] == d->N[1] && d->N[1] == d->N[2]))
    return gf_set_err(ctx, GF_ERR_ARG, "device GGFX maps need 3-D regular isotropic boxes");
  if (!devmaps && (!d->ggfx_offs || !d->ggfx_members))
    return gf_set_err(ctx, GF_ERR_ARG, "missing GGFX maps");
  GF_HIP(ctx, hipSetDevice(ctx->device));
  gf_mesh *m = new gf_mesh_s();
  m->ctx = ctx;
  m->dim = d->dim;
  m->gtype = d->gtype;
  m->nelems = d->nelems;
  m->nn = 1;
  for (int k = 0; k < 3; k++) m->N[k] = (k < d->dim) ? d->N[k] : 1;
  for (int k = 0; k < d->dim; k++) {
    if (m->N[k] < 2 || m->N[k] > 64 || !d->D[k] || !d->W[k]) {
      delete m;
      return gf_set_err(ctx, GF_ERR_ARG, "bad basis");
    }
    m->nn *= (size_t)m->N[k];
  }
  m->ndof = m->nelems * m->nn;
  bool iso = true;
  for (int k = 1; k < d->dim; k++) iso = iso && m->N[k] == m->N[0];
  int nmax = d->dim == 3 ? 10 : 12;
  m->Niso = (iso && m->N[0] <= nmax) ? m->N[0] : 0;
  // the constant-D pencil kernels serve regular elements only
  if (m->dim == 3 && m->Niso >= 2 && m->Niso <= 10 && d->gtype == GF_GE_REGULAR) {
    bool same = true;
    for (int k = 0; k < 3 && same; k++)
      for (int l = 0; l < m->Niso && same; l++)
        for (int q = 0; q < m->Niso && same; q++)
          same = d->D[k][l + q * m->Niso] == gll_const(m->Niso, l, q);
    m->dconst = same;
  }
  int rc = GF_OK;
#define UP(dst, src, n)                      \
  do {                                       \
    rc = upload(ctx, &(dst), (src), (n));    \
    if (rc) { gf_mesh_destroy(m); return rc; } \
  } while (0)
  for (int k = 0; k < d->dim; k++) {
    UP(m->D[k], d->D[k], (size_t)m->N[k] * m->N[k]);
    UP(m->W[k], d->W[k], (size_t)m->N[k]);
  }
  if (m->Niso) {
    std::vector<float> Df((size_t)m->Niso * m->Niso), Wf(m->Niso);
    for (size_t q = 0; q < Df.size(); q++) Df[q] = (float)d->D[0][q];
    for (int q = 0; q < m->Niso; q++) Wf[q] = (float)d->W[0][q];
    UP(m->D32, Df.data(), Df.size());
    UP(m->W32, Wf.data(), Wf.size());
    int Nn = m->Niso;
    std::vector<double> Rm((size_t)Nn * Nn), Cm((size_t)Nn * Nn);
    std::vector<float> Rf((size_t)Nn * Nn), Cf((size_t)Nn * Nn);
    for (int l = 0; l < Nn; l++)
      for (int q = 0; q < Nn; q++) {
        Rm[l * Nn + q] = d->D[0][l + q * Nn];  // D(l,q)
        Cm[l * Nn + q] = d->D[0][q + l * Nn];  // D(q,l)
        Rf[l * Nn + q] = (float)Rm[l * Nn + q];
        Cf[l * Nn + q] = (float)Cm[l * Nn + q];
      }
    for (int c = 0; c < 3; c++) {
      UP(m->pDR[c], Rm.data(), Rm.size());
      UP(m->pDC[c], Cm.data(), Cm.size());
      UP(m->pDR32[c], Rf.data(), Rf.size());
      UP(m->pDC32[c], Cf.data(), Cf.size());
    }
  }
  size_t n = m->ndof;
  if (d->gtype == GF_GE_REGULAR) {
    if (!d->elem_dxidx || !d->elem_jac) {
      gf_mesh_destroy(m);
      return gf_set_err(ctx, GF_ERR_ARG, "regular mesh needs elem_dxidx and elem_jac");
    }
    UP(m->edx, d->elem_dxidx, m->nelems * d->dim);
    UP(m->ejac, d->elem_jac, m->nelems);
    std::vector<float> a(m->nelems * d->dim), b(m->nelems);
    for (size_t q = 0; q < a.size(); q++) a[q] = (float)d->elem_dxidx[q];
    for (size_t q = 0; q < b.size(); q++) b[q] = (float)d->elem_jac[q];
    UP(m->edx32, a.data(), a.size());
    UP(m->ejac32, b.data(), b.size());
  } else {
    // Cartesian coordinates per node: 3 on an embedded surface, else dim
    const int nxy = d->gtype == GF_GE_2DEMBEDDED ? 3 : d->dim;
    if (!d->node_jac) {
      gf_mesh_destroy(m);
      return gf_set_err(ctx, GF_ERR_ARG, "embedded or deformed mesh needs node metrics");
    }
    for (int i = 0; i < d->dim; i++)
      for (int j = 0; j < nxy; j++) {
        if (!d->node_dxidx[i * 3 + j]) {
          gf_mesh_destroy(m);
          return gf_set_err(ctx, GF_ERR_ARG, "node metrics need dXi_i/dX_j for i<dim, j<nxy");
        }
        UP(m->ndx[i * 3 + j], d->node_dxidx[i * 3 + j], n);
      }
    UP(m->njac, d->node_jac, n);
    // Metric G(i,j) = sum_{k<nxy} dXidX(i,k) dXidX(j,k) W0 W1 Jac (def_init
    // 2-D, src/cdg/sem/ghelmholtz.ipp:476-491), built on the host in that order.
    std::vector<double> G[6];
    int Ni = m->N[0], Nj = m->N[1];
    for (int c = 0; c < 3; c++) G[c].assign(n, 0.0);
    const int ii[3] = {0, 0, 1}, jj[3] = {0, 1, 1};
    for (int c = 0; c < (d->dim == 2 ? 3 : 0); c++) {
      std::fill(G[c].begin(), G[c].end(), 0.0);
      size_t q = 0;
      for (size_t e = 0; e < m->nelems; e++)
        for (int mm = 0; mm < Nj; mm++)
          for (int l = 0; l < Ni; l++, q++)
            for (int k = 0; k < nxy; k++)
              G[c][q] += d->node_dxidx[ii[c] * 3 + k][q] * d->node_dxidx[jj[c] * 3 + k][q] *
                         d->W[0][l] * d->W[1][mm] * d->node_jac[q];
      UP(m->G[c], G[c].data(), n);
    }
    if (d->dim == 3) {
      // def_init 3-D (ghelmholtz.ipp:494-514): both loops run over all (i, j)
      // while G(i,j) and G(j,i) share one vector, so each off-diagonal vector
      // receives its terms twice, in this order of passes.  Slots:
      // 0 G00, 1 G01, 2 G11, 3 G02, 4 G12, 5 G22.
      auto slot = [](int i, int j) {
        const int a = i < j ? i : j, b = i < j ? j : i;
        return a == b ? (a == 0 ? 0 : a == 1 ? 2 : 5) : (b == 1 ? 1 : a == 0 ? 3 : 4);
      };
      const int Nk = m->N[2];
      for (int c = 0; c < 6; c++) G[c].assign(n, 0.0);
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          std::vector<double> &Gv = G[slot(i, j)];
          size_t q = 0;
          for (size_t e = 0; e < m->nelems; e++)
            for (int pp = 0; pp < Nk; pp++)
              for (int mm = 0; mm < Nj; mm++)
                for (int l = 0; l < Ni; l++, q++)
                  for (int k = 0; k < 3; k++)
                    Gv[q] += d->node_dxidx[i * 3 + k][q] * d->node_dxidx[j * 3 + k][q] *
                             d->W[0][l] * d->W[1][mm] * d->W[2][pp] * d->node_jac[q];
        }
      for (int c = 0; c < 6; c++) UP(m->G[c], G[c].data(), n);
    }
  }
  for (int k = 0; k < 3; k++)
    if (d->xyz[k]) UP(m->xyz[k], d->xyz[k], n);
  // mask
  m->has_mask = d->mask != nullptr;
  {
    std::vector<double> mk(n, 1.0);
    if (d->mask) std::copy(d->mask, d->mask + n, mk.begin());
    UP(m->mask, mk.data(), n);
    std::vector<float> mf(n);
    for (size_t q = 0; q < n; q++) mf[q] = (float)mk[q];
    UP(m->mask32, mf.data(), n);
  }
  // DSS groups with more than one member, each stored once, in order of
  // their smallest member (which is the node at which GGFX::init formed it).
  if (devmaps) {
    if (boxunion) {
      std::vector<int> ne(3 * (size_t)d->box_count), per(3 * (size_t)d->box_count);
      for (int b = 0; b < d->box_count; b++)
        for (int c = 0; c < 3; c++) {
          ne[3 * b + c] = d->box_list[6 * b + c];
          per[3 * b + c] = d->box_list[6 * b + 3 + c];
        }
      rc = gf_box_groups_device(m, d->box_count, ne.data(), per.data());
    } else {
      rc = gf_box_groups_device(m, 1, d->box_ne, d->box_periodic);
    }
    if (rc) {
      gf_mesh_destroy(m);
      return rc;
    }
  } else {
    std::vector<int> goff(1, 0), gmem;
    int maxg = 1;
    for (size_t q = 0; q < n; q++) {
      int64_t b = d->ggfx_offs[q], e = d->ggfx_offs[q + 1];
      if (e - b < 1 || d->ggfx_members[b] < 0 || (size_t)d->ggfx_members[e - 1] >= n) {
        gf_mesh_destroy(m);
        return gf_set_err(ctx, GF_ERR_ARG, "malformed GGFX map");
      }
      if (e - b > 1 && (size_t)d->ggfx_members[b] == q) {
        for (int64_t s = b; s < e; s++) gmem.push_back((int)d->ggfx_members[s]);
        goff.push_back((int)gmem.size());
        maxg = std::max<int>(maxg, (int)(e - b));
      }
    }
    m->ngroups = (int)goff.size() - 1;
    m->maxgrp = maxg;
    UP(m->goff, goff.data(), goff.size());
    UP(m->gmem, gmem.data(), std::max<size_t>(gmem.size(), 1));
  }
  // Structured box: enable gather assembly only if every node's GGFX group
  // is exactly the group the box structure implies.
  if (d->box_ne[0] > 0 && !boxunion && d->dim == 3 && d->gtype == GF_GE_REGULAR && m->Niso) {
    gf::BoxTopo t;
    size_t ne_tot = 1;
    for (int k = 0; k < 3; k++) {
      t.ne[k] = d->box_ne[k];
      t.per[k] = d->box_periodic[k] ? 1 : 0;
      ne_tot *= (size_t)std::max(d->box_ne[k], 0);
    }
    for (int f = 0; f < 6; f++) t.fmask[f] = 0;
    bool ok = ne_tot == m->nelems;
    const int Nn = m->Niso;
    // face mask types, read off the mask at the middle of each face (checked
    // against every node below)
    if (ok && d->mask) {
      for (int f = 0; f < 6; f++) {
        const int dd = f / 2, side = f % 2;
        int ec[3], nc[3];
        for (int c = 0; c < 3; c++) {
          ec[c] = c == dd ? (side ? t.ne[c] - 1 : 0) : t.ne[c] / 2;
          nc[c] = c == dd ? (side ? Nn - 1 : 0) : Nn / 2;
        }
        const size_t e = (size_t)ec[0] + (size_t)t.ne[0] * ((size_t)ec[1] + (size_t)t.ne[1] * ec[2]);
        const size_t g = e * m->nn + (size_t)nc[2] * Nn * Nn + (size_t)nc[1] * Nn + nc[0];
        t.fmask[f] = d->mask[g] == 0.0 ? 1 : 0;
      }
    }
    bool mask_ok = ok;
    // device-built maps are the box's groups by construction; only the mask
    // is checked node by node
    for (size_t q = 0; ok && q < n && !(devmaps && !mask_ok); q++) {
      const size_t e = q / m->nn, r = q - e * m->nn;
      const int i = (int)(r % Nn), j = (int)((r / Nn) % Nn), k = (int)(r / ((size_t)Nn * Nn));
      // the construction the kernels use: per-element partner offsets
      int64_t g[8];
      int ec[3];
      gf::box_elem_coords(t, e, ec);
      const gf::BoxElemDeltas de = gf::box_elem_deltas(t, Nn, ec);
      if (!devmaps) {
        const int c = gf::box_members<int64_t>(de, Nn, (int64_t)q, i, j, k, g);
        const int64_t b = d->ggfx_offs[q], eo = d->ggfx_offs[q + 1];
        ok = (eo - b) == c;
        for (int s2 = 0; ok && s2 < c; s2++) ok = d->ggfx_members[b + s2] == g[s2];
      }
      if (mask_ok) {
        const double mv = d->mask ? d->mask[q] : 1.0;
        mask_ok = (mv == 0.0 || mv == 1.0) && ((mv == 0.0) == gf::box_masked(t, Nn, ec, i, j, k));
      }
    }
    m->box = t;
    m->has_box = ok;
    m->has_box_mask = ok && mask_ok;
  }
  // the LDS-halo gather of the N = 8 stage kernels (gf::box_halo8_ok)
  m->halo8 = m->has_box && m->Niso == 8 && gf::box_halo8_ok(m->box);
  m->red_cap = gf_red_need(m->nelems);
  GF_HIP(ctx, hipMalloc((void **)&m->red, m->red_cap * sizeof(double)));
  GF_HIP(ctx, hipMalloc((void **)&m->scal, 64 * sizeof(double)));
  GF_HIP(ctx, hipHostMalloc((void **)&m->hscal, 64 * sizeof(double),
                            hipHostMallocMapped | hipHostMallocCoherent));
  GF_HIP(ctx, hipHostGetDevicePointer((void **)&m->hscal_dev, m->hscal, 0));
  for (int k = 0; k < 2; k++)
    GF_HIP(ctx, hipEventCreateWithFlags(&m->cg_ev[k], hipEventDisableTiming));
  GF_HIP(ctx, hipMalloc((void **)&m->mass, n * sizeof(double)));
  GF_HIP(ctx, hipMalloc((void **)&m->imass, n * sizeof(double)));
  GF_HIP(ctx, hipMalloc((void **)&m->imult, n * sizeof(double)));
  GF_HIP(ctx, hipMalloc((void **)&m->imult32, n * sizeof(float)));
  hipStream_t s = ctx->stream;
  // mass (GMass::init), then inverse via DSS-Smooth (gmass.ipp:75-82)
  if (d->gtype == GF_GE_REGULAR && m->Niso) {
    rc = GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_mass_init, m);
  } else {
    hipLaunchKernelGGL(k_mass_init_generic, dim3(nblocks(n)), dim3(kThreads), 0, s, m->mass,
                       m->W[0], m->W[1], m->dim == 3 ? m->W[2] : m->W[1],
                       d->gtype == GF_GE_REGULAR ? m->ejac : m->njac,
                       d->gtype == GF_GE_REGULAR ? 0 : 1, m->N[0], m->N[1], m->N[2], m->dim, n);
  }
  if (rc) { gf_mesh_destroy(m); return rc; }
  GF_HIP(ctx, hipMemcpyAsync(m->imass, m->mass, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  rc = gf_dss_t(m, GF_DSS_SMOOTH, m->imass);
  if (rc) { gf_mesh_destroy(m); return rc; }
  hipLaunchKernelGGL(k_recip, dim3(nblocks(n)), dim3(kThreads), 0, s, m->imass, n);
  if (d->gtype != GF_GE_REGULAR) {
    // FP32 copies of the node metrics and the mass for the mixed-precision CG
    for (int c = 0; c < 7; c++) {
      const double *src = c < 6 ? m->G[c] : m->mass;
      if (!src) continue;
      float *dst = nullptr;
      GF_HIP(ctx, hipMalloc((void **)&dst, n * sizeof(float)));
      if (c < 6) m->G32[c] = dst;
      else m->mass32 = dst;
      hipLaunchKernelGGL(k_narrow_f32, dim3(nblocks(n)), dim3(kThreads), 0, s, dst, src, n);
    }
  }
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kThreads), 0, s, m->imult, 1.0, n);
  if (m->ngroups)
    hipLaunchKernelGGL(k_imult_groups, dim3(nblocks((size_t)m->ngroups)), dim3(kThreads), 0, s,
                       m->imult, m->goff, m->gmem, m->ngroups);
  GF_HIP(ctx, hipStreamSynchronize(s));
  {
    std::vector<double> im(n);
    GF_HIP(ctx, hipMemcpy(im.data(), m->imult, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> imf(n);
    for (size_t q = 0; q < n; q++) imf[q] = (float)im[q];
    GF_HIP(ctx, hipMemcpy(m->imult32, imf.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  if (m->has_box && m->dconst) {
    // Jacobi-diagonal table of the box CG (filled per solve)
    GF_HIP(ctx, hipMalloc((void **)&m->dtab, m->nn * sizeof(double)));
    GF_HIP(ctx, hipMalloc((void **)&m->dtab32, m->nn * sizeof(float)));
    GF_HIP(ctx, hipMalloc((void **)&m->deq, m->nelems));
    m->dref = (size_t)(m->box.ne[0] / 2) +
              (size_t)m->box.ne[0] * ((size_t)(m->box.ne[1] / 2) +
                                      (size_t)m->box.ne[1] * (size_t)(m->box.ne[2] / 2));
  }
  if (d->gtype == GF_GE_REGULAR && m->dim == 3 && m->Niso && m->nelems > 1) {
    // inverse mass identical at every element (bit for bit): the RK stages
    // read one element's values instead of the whole array
    std::vector<double> im(n);
    GF_HIP(ctx, hipMemcpy(im.data(), m->imass, n * sizeof(double), hipMemcpyDeviceToHost));
    const size_t nn = m->nn;
    bool same = true;
    for (size_t q = nn; same && q < n; q++) same = im[q] == im[q % nn];
    if (same) {
      GF_HIP(ctx, hipMalloc((void **)&m->imass_el, nn * sizeof(double)));
      GF_HIP(ctx, hipMemcpy(m->imass_el, im.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    }
  }
#undef UP
  *out = m;
  return GF_OK;
}

int gf_mesh_flags(const gf_mesh *m) {
  if (!m) return 0;
  return (m->has_box ? GF_MESH_BOX_GATHER : 0) | (m->has_box_mask ? GF_MESH_BOX_MASK : 0) |
         (m->dconst ? GF_MESH_CONST_D : 0) | (m->halo8 ? GF_MESH_BOX_HALO : 0) |
         (m->imass_el ? GF_MESH_IMASS_ELEM : 0);
}

int gf_mesh_destroy(gf_mesh *m) {
  if (!m) return GF_OK;
  auto F = [](void *p) {
    if (p) (void)hipFree(p);
  };
  for (int k = 0; k < 3; k++) {
    F(m->D[k]);
    F(m->W[k]);
    F(m->xyz[k]);
  }
  for (int k = 0; k < 6; k++) F(m->G[k]);
  for (int k = 0; k < 9; k++) F(m->ndx[k]);
  for (int k = 0; k < 3; k++) {
    F(m->pDR[k]);
    F(m->pDC[k]);
    F(m->pDR32[k]);
    F(m->pDC32[k]);
  }
  F(m->D32); F(m->W32); F(m->edx); F(m->ejac); F(m->edx32); F(m->ejac32); F(m->njac);
  F(m->mask); F(m->mass); F(m->imass); F(m->imass_el); F(m->dtab); F(m->dtab32); F(m->deq);
  F(m->imult); F(m->mask32);
  F(m->imult32);
  F(m->goff); F(m->gmem); F(m->red); F(m->scal);
  for (int k = 0; k < 6; k++) F(m->G32[k]);
  F(m->mass32);
  if (m->hscal) (void)hipHostFree(m->hscal);
  F(m->seg_eoff); F(m->seg_boff); F(m->sscal); F(m->dl_I); F(m->dl_wq);
  if (m->hsscal) (void)hipHostFree(m->hsscal);
  for (int k = 0; k < 2; k++)
    if (m->cg_ev[k]) (void)hipEventDestroy(m->cg_ev[k]);
  for (size_t k = 0; k < m->ws.size(); k++) F(m->ws[k]);
  delete m;
  return GF_OK;
}

size_t gf_mesh_ndof(const gf_mesh *m) { return m ? m->ndof : 0; }

int gf_mesh_get(gf_mesh *m, const char *which, double *host_out) {
  if (!m || !which || !host_out) return GF_ERR_ARG;
  std::string w(which);
  const double *src = nullptr;
  if (w == "mass") src = m->mass;
  else if (w == "imass") src = m->imass;
  else if (w == "imult") src = m->imult;
  else if (w == "mask") src = m->mask;
  if (!src) return gf_set_err(m->ctx, GF_ERR_ARG, "unknown mesh array");
  GF_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  GF_HIP(m->ctx, hipMemcpy(host_out, src, m->ndof * sizeof(double), hipMemcpyDeviceToHost));
  return GF_OK;
}

// ===========================================================================
// Operators
// ===========================================================================
int gf_grefderiv(gf_mesh *m, int idir, int dotrans, const double *u, double *du) {
  if (!m || !u || !du) return GF_ERR_ARG;
  if (idir < 1 || idir > m->dim) {
    if (m->dim == 2 && idir == 3) {  // grefderiv 2-D idir 3 is zero (ggrid.ipp:2133-2137)
      GF_HIP(m->ctx, hipMemsetAsync(du, 0, m->ndof * sizeof(double), m->ctx->stream));
      return GF_OK;
    }
    return gf_set_err(m->ctx, GF_ERR_ARG, "idir out of range");
  }
  return gf_grefderiv_t(m, idir, dotrans, u, du, nullptr);
}

int gf_grefderivs(gf_mesh *m, int dotrans, const double *u, double *const *du) {
  if (!m || !u || !du) return GF_ERR_ARG;
  if (!dotrans && m->Niso) {
    for (int k = 0; k < m->dim; k++)
      if (!du[k]) return GF_ERR_ARG;
    return GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_grad, m, u, du[0], du[1],
                         m->dim == 3 ? du[2] : nullptr, (const double *)nullptr);
  }
  for (int k = 0; k < m->dim; k++) {
    int rc = gf_grefderiv(m, k + 1, dotrans, u, du[k]);
    if (rc) return rc;
  }
  return GF_OK;
}

int gf_grefdiv(gf_mesh *m, int dotrans, const double *const *u, double *divu) {
  if (!m || !u || !divu) return GF_ERR_ARG;
  if (!m->Niso) return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "grefdiv needs isotropic order");
  return GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_div, m, dotrans, u[0], u[1],
                       m->dim == 3 ? u[2] : nullptr, divu);
}

int gf_deriv(gf_mesh *m, int idir, const double *u, double *du) {
  if (!m || !u || !du) return GF_ERR_ARG;
  if (m->gtype == GF_GE_REGULAR) {
    if (idir < 1 || idir > m->dim) return gf_set_err(m->ctx, GF_ERR_ARG, "idir out of range");
    return gf_grefderiv_t(m, idir, 0, u, du, m->edx);
  }
  // embedded or deformed (GGrid::deriv, ggrid.ipp:1179-1189):
  // du = (D1 u) dXidX(0,idir-1); du += (D_j u) dXidX(j-1,idir-1), j = 2..dim
  if (idir < 1 || idir > (m->gtype == GF_GE_2DEMBEDDED ? 3 : m->dim))
    return gf_set_err(m->ctx, GF_ERR_ARG, "idir out of range");
  double *t = (double *)gf_ws(m, WS_TMP0, m->ndof * sizeof(double));
  double *t1 = (double *)gf_ws(m, WS_TMP1, m->ndof * sizeof(double));
  if (!t || !t1) return gf_set_err(m->ctx, GF_ERR_ALLOC, "workspace");
  int rc = gf_grefderiv_t(m, 1, 0, u, du, nullptr);
  if (rc) return rc;
  rc = gf_mul_t(m, du, m->ndx[0 * 3 + idir - 1], du);
  if (rc) return rc;
  for (int j = 1; j < m->dim; j++) {
    rc = gf_grefderiv_t(m, j + 1, 0, u, t, nullptr);
    if (rc) return rc;
    rc = gf_mul_t(m, t, m->ndx[j * 3 + idir - 1], t);
    if (rc) return rc;
    rc = gf_axpy_t(m, 1.0, t, du);
    if (rc) return rc;
  }
  return GF_OK;
}

int gf_grad(gf_mesh *m, const double *u, double *const *du) {
  if (!m || !u || !du) return GF_ERR_ARG;
  if (m->gtype == GF_GE_REGULAR && m->Niso) {
    for (int k = 0; k < m->dim; k++)
      if (!du[k]) return GF_ERR_ARG;
    return GF_DISPATCH_N(m->dim, m->Niso, gfl::launch_grad, m, u, du[0], du[1],
                         m->dim == 3 ? du[2] : nullptr, (const double *)m->edx);
  }
  int nd = m->gtype == GF_GE_2DEMBEDDED ? 3 : m->dim;
  for (int k = 0; k < nd; k++) {
    int rc = gf_deriv(m, k + 1, u, du[k]);
    if (rc) return rc;
  }
  return GF_OK;
}

int gf_mass_apply(gf_mesh *m, const double *in, double *out) {
  if (!m || !in || !out) return GF_ERR_ARG;
  return gf_mul_t(m, m->mass, in, out);
}

int gf_imass_apply(gf_mesh *m, const double *in, double *out) {
  if (!m || !in || !out) return GF_ERR_ARG;
  return gf_mul_t(m, m->imass, in, out);
}

int gf_integrate(gf_mesh *m, const double *u, double *result) {
  if (!m || !u || !result) return GF_ERR_ARG;
  return gf_dot_t(m, m->mass, u, m->ndof, result);
}

// GGrid::dealias (ggrid.ipp:2424-2503): one workgroup per element; v1 and v2
// are interpolated to the order-pq GLL grid direction by direction in LDS
// (x, then y, then z), multiplied pointwise and by the q tensor weights
// ((Wx Wy) Wz), projected back with the transposed interpolation in the same
// direction order and multiplied by iWp = Jac * imass (init_qdealias,
// ggrid.ipp:2402-2407).  I is column-major Q x N: I[a + Q i] = h_i(xq_a).
namespace {
__device__ void dl_pass(const double *__restrict__ in, double *__restrict__ out, int n0, int n1,
                        int n2, int d, int lout, const double *__restrict__ I, int Q, bool back) {
  const int lin = d == 0 ? n0 : (d == 1 ? n1 : n2);
  const int m0 = d == 0 ? lout : n0, m1 = d == 1 ? lout : n1, m2 = d == 2 ? lout : n2;
  const int tot = m0 * m1 * m2;
  for (int o = threadIdx.x; o < tot; o += blockDim.x) {
    const int a = o % m0, b = (o / m0) % m1, c = o / (m0 * m1);
    const int oi = d == 0 ? a : (d == 1 ? b : c);
    double acc = 0.0;
    for (int k = 0; k < lin; k++) {
      const int ia = d == 0 ? k : a, ib = d == 1 ? k : b, ic = d == 2 ? k : c;
      const double op = back ? I[k + Q * oi] : I[oi + Q * k];
      acc = fma(op, in[ia + n0 * (ib + n1 * ic)], acc);
    }
    out[o] = acc;
  }
  __syncthreads();
}

__global__ void k_dealias(const double *__restrict__ v1, const double *__restrict__ v2,
                          double *__restrict__ prod, const double *__restrict__ I,
                          const double *__restrict__ wq, const double *__restrict__ ejac,
                          const double *__restrict__ imass, int dim, int N, int Q) {
  extern __shared__ double sh[];
  const int z = dim == 3 ? 1 : 0;
  const int NN = z ? N * N * N : N * N, QQ = z ? Q * Q * Q : Q * Q;
  double *sI = sh, *sw = sI + Q * N, *sv = sw + Q, *a1 = sv + NN, *a2 = a1 + QQ, *b1 = a2 + QQ,
         *b2 = b1 + QQ;
  const size_t e = blockIdx.x;
  for (int q = threadIdx.x; q < Q * N; q += blockDim.x) sI[q] = I[q];
  for (int q = threadIdx.x; q < Q; q += blockDim.x) sw[q] = wq[q];
  for (int f = 0; f < 2; f++) {
    const double *v = (f == 0 ? v1 : v2) + e * NN;
    for (int q = threadIdx.x; q < NN; q += blockDim.x) sv[q] = v[q];
    __syncthreads();
    double *dst = f == 0 ? b1 : b2;
    dl_pass(sv, a1, N, N, z ? N : 1, 0, Q, sI, Q, false);
    if (z) {
      dl_pass(a1, a2, Q, N, N, 1, Q, sI, Q, false);
      dl_pass(a2, dst, Q, Q, N, 2, Q, sI, Q, false);
    } else {
      dl_pass(a1, dst, Q, N, 1, 1, Q, sI, Q, false);
    }
  }
  for (int o = threadIdx.x; o < QQ; o += blockDim.x) {
    const int a = o % Q, b = (o / Q) % Q, c = o / (Q * Q);
    double w = sw[a] * sw[b];
    if (z) w = w * sw[c];
    const double t = b1[o] * b2[o];
    b1[o] = t * w;
  }
  __syncthreads();
  dl_pass(b1, a1, Q, Q, z ? Q : 1, 0, N, sI, Q, true);
  double *res = z ? b2 : a2;
  if (z) {
    dl_pass(a1, a2, N, Q, Q, 1, N, sI, Q, true);
    dl_pass(a2, res, N, N, Q, 2, N, sI, Q, true);
  } else {
    dl_pass(a1, res, N, Q, 1, 1, N, sI, Q, true);
  }
  const double jac = ejac[e];
  for (int q = threadIdx.x; q < NN; q += blockDim.x) {
    const size_t g = e * NN + q;
    prod[g] = res[q] * (jac * imass[g]);
  }
}
}  // namespace

int gf_dealias(gf_mesh *m, int pq, const double *v1, const double *v2, double *prod) {
  if (!m || !v1 || !v2 || !prod) return GF_ERR_ARG;
  if (m->gtype != GF_GE_REGULAR || !m->Niso)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "dealias: regular meshes of constant p");
  const int N = m->Niso, p = N - 1, Q = pq + 1;
  if (pq < p || pq > 31) return gf_set_err(m->ctx, GF_ERR_ARG, "dealias: need p <= pq <= 31");
  const int z = m->dim == 3 ? 1 : 0;
  const size_t NN = z ? (size_t)N * N * N : (size_t)N * N, QQ = z ? (size_t)Q * Q * Q : (size_t)Q * Q;
  const size_t lds = ((size_t)Q * N + Q + NN + 4 * QQ) * sizeof(double);
  if (lds > 160 * 1024)
    return gf_set_err(m->ctx, GF_ERR_UNSUPPORTED, "dealias: element does not fit in LDS");
  if (m->dealias_pq != pq) {
    std::vector<double> I((size_t)Q * N), wq(Q);
    if (gf_dealias_matrix(p, pq, I.data(), wq.data()) != GF_OK)
      return gf_set_err(m->ctx, GF_ERR_ARG, "dealias: basis");
    if (m->dl_I) (void)hipFree(m->dl_I);
    if (m->dl_wq) (void)hipFree(m->dl_wq);
    m->dl_I = m->dl_wq = nullptr;
    m->dealias_pq = -1;
    GF_HIP(m->ctx, hipMalloc((void **)&m->dl_I, I.size() * sizeof(double)));
    GF_HIP(m->ctx, hipMalloc((void **)&m->dl_wq, wq.size() * sizeof(double)));
    GF_HIP(m->ctx, hipMemcpy(m->dl_I, I.data(), I.size() * sizeof(double), hipMemcpyHostToDevice));
    GF_HIP(m->ctx, hipMemcpy(m->dl_wq, wq.data(), wq.size() * sizeof(double), hipMemcpyHostToDevice));
    m->dealias_pq = pq;
  }
  if (lds > 64 * 1024)
    GF_HIP(m->ctx, hipFuncSetAttribute((const void *)k_dealias,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_dealias, dim3(m->nelems), dim3(kThreads), lds, m->ctx->stream, v1, v2, prod,
                     m->dl_I, m->dl_wq, m->ejac, m->imass, m->dim, N, Q);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// gf_integrate's kernels and block count with the partial sums in this
// device's scratch array; the result stays on the device; asynchronous.
int gf_integrate_noalloc(gf_mesh *m, const double *u, double *result) {
  if (!m || !u || !result) return GF_ERR_ARG;
  double *part = m->ctx->scratch;
  if (!part) return gf_set_err(m->ctx, GF_ERR_ARG, "integrate_noalloc: context has no scratch");
  const int nb = std::min(nblocks(m->ndof), 1024);
  if (nb > kNoAllocParts)
    return gf_set_err(m->ctx, GF_ERR_ARG, "integrate_noalloc: more partial sums than scratch slots");
  hipLaunchKernelGGL(k_partial_dot, dim3(nb), dim3(kThreads), 0, m->ctx->stream, m->mass, u, part,
                     m->ndof);
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kThreads), 0, m->ctx->stream, part, nb, result);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

int gf_helmholtz_apply(gf_mesh *m, double p_scalar, const double *p_field, double q_scalar,
                       const double *q_field, const double *u, double *uo) {
  if (!m || !u || !uo) return GF_ERR_ARG;
  return gf_helm_t(m, p_scalar, p_field, q_field, q_scalar, u, uo);
}

int gf_dss(gf_mesh *m, int op, double *u) {
  if (!m || !u) return GF_ERR_ARG;
  if (op < GF_DSS_SUM || op > GF_DSS_MIN) return gf_set_err(m->ctx, GF_ERR_ARG, "bad DSS op");
  return gf_dss_t(m, op, u);
}

int gf_imult(gf_mesh *m, double *imult) {
  if (!m || !imult) return GF_ERR_ARG;
  GF_HIP(m->ctx, hipMemcpyAsync(imult, m->imult, m->ndof * sizeof(double),
                                hipMemcpyDeviceToDevice, m->ctx->stream));
  return GF_OK;
}

}  // extern "C"
