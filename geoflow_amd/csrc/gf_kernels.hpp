// gf_kernels.hpp -- device building blocks for the SEM operator kernels.
//
// Every element-local kernel stages whole elements of a field into LDS
// (elements are contiguous in HBM, so the staging loads are fully coalesced)
// and evaluates the 1-D GLL contractions from LDS.  A block of 256 threads
// (four wave64s) owns EPB = max(1, 256 / N^d) elements.
#pragma once

#include <hip/hip_runtime.h>

#include "gf_internal.hpp"

namespace gf {

constexpr int kThreads = 256;

template <int DIM, int N>
struct Shape {
  static constexpr int NN = DIM == 3 ? N * N * N : N * N;
  static constexpr int EPB = NN >= kThreads ? 1 : kThreads / NN;
  static constexpr int NODES = EPB * NN;
};

// Node coordinates inside an element from the in-element index r.
template <int DIM, int N>
__device__ __forceinline__ void node_ijk(int r, int &i, int &j, int &k) {
  i = r % N;
  j = (r / N) % N;
  k = DIM == 3 ? r / (N * N) : 0;
}

// Contraction along direction dir (0,1,2) of an element staged at su:
//   y = sum_m Dop(o, m) u(..m..),  Dop = D (TRANS=false) or D^T (TRANS=true),
// with D column-major in sD: D(o,m) = sD[o + m*N].  Terms are accumulated in
// ascending m, the order of the reference's dmxm (cff_blas.F, dmxmp.F).
template <int DIM, int N, bool TRANS, typename T>
__device__ __forceinline__ T contract(const T *__restrict__ su, const T *__restrict__ sD, int dir,
                                      int i, int j, int k) {
  T acc = T(0);
  if (dir == 0) {
    const T *row = su + (DIM == 3 ? k * N * N : 0) + j * N;
#pragma unroll
    for (int m = 0; m < N; m++) {
      T d = TRANS ? sD[m + i * N] : sD[i + m * N];
      acc = fma(d, row[m], acc);
    }
  } else if (dir == 1) {
    const T *col = su + (DIM == 3 ? k * N * N : 0) + i;
#pragma unroll
    for (int m = 0; m < N; m++) {
      T d = TRANS ? sD[m + j * N] : sD[j + m * N];
      acc = fma(d, col[m * N], acc);
    }
  } else {
    const T *col = su + j * N + i;
#pragma unroll
    for (int m = 0; m < N; m++) {
      T d = TRANS ? sD[m + k * N] : sD[k + m * N];
      acc = fma(d, col[m * N * N], acc);
    }
  }
  return acc;
}

// Stage elements [e0, e0+nel) of field u into LDS buffer s (contiguous).
template <typename T>
__device__ __forceinline__ void stage(T *__restrict__ s, const T *__restrict__ u, size_t base,
                                      int count) {
  for (int q = threadIdx.x; q < count; q += blockDim.x) s[q] = u[base + q];
}

template <typename T>
__device__ __forceinline__ void stage_D(T *__restrict__ sD, const T *__restrict__ D, int N) {
  for (int q = threadIdx.x; q < N * N; q += blockDim.x) sD[q] = D[q];
}

// Lumped mass at an element node: ((W[k] W[j]) W[i]) Jac in 3-D, (W[j] W[i]) Jac
// in 2-D -- GMass::init3d/init2d followed by pointProd(Jac)
// (src/cdg/sem/gmass.ipp:134-212).
template <int DIM, typename T>
__device__ __forceinline__ T node_mass(const T *__restrict__ sW, int i, int j, int k, T jac) {
  T w = DIM == 3 ? sW[k] * sW[j] * sW[i] : sW[j] * sW[i];
  return w * jac;
}

// ---- block reductions (wave64) -------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
// Sums (MAXMODE=false) or maxes over the block; result valid in thread 0.
template <bool MAXMODE>
__device__ __forceinline__ double block_reduce(double v, double *sred) {
  v = MAXMODE ? wave_max(v) : wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sred[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    r = sred[0];
    for (int w = 1; w < nw; w++) r = MAXMODE ? fmax(r, sred[w]) : r + sred[w];
  }
  return r;
}

}  // namespace gf

// Dispatch an isotropic-N template functor F<DIM,N>::run(args...) at runtime.
#define GF_DISPATCH_N(DIMV, NV, FN, ...)                         \
  [&]() -> int {                                                 \
    if ((DIMV) == 3) {                                           \
      switch (NV) {                                              \
        case 2: return FN<3, 2>(__VA_ARGS__);                    \
        case 3: return FN<3, 3>(__VA_ARGS__);                    \
        case 4: return FN<3, 4>(__VA_ARGS__);                    \
        case 5: return FN<3, 5>(__VA_ARGS__);                    \
        case 6: return FN<3, 6>(__VA_ARGS__);                    \
        case 7: return FN<3, 7>(__VA_ARGS__);                    \
        case 8: return FN<3, 8>(__VA_ARGS__);                    \
        case 9: return FN<3, 9>(__VA_ARGS__);                    \
        case 10: return FN<3, 10>(__VA_ARGS__);                  \
        default: return GF_ERR_UNSUPPORTED;                      \
      }                                                          \
    } else {                                                     \
      switch (NV) {                                              \
        case 2: return FN<2, 2>(__VA_ARGS__);                    \
        case 3: return FN<2, 3>(__VA_ARGS__);                    \
        case 4: return FN<2, 4>(__VA_ARGS__);                    \
        case 5: return FN<2, 5>(__VA_ARGS__);                    \
        case 6: return FN<2, 6>(__VA_ARGS__);                    \
        case 7: return FN<2, 7>(__VA_ARGS__);                    \
        case 8: return FN<2, 8>(__VA_ARGS__);                    \
        case 9: return FN<2, 9>(__VA_ARGS__);                    \
        case 10: return FN<2, 10>(__VA_ARGS__);                  \
        case 11: return FN<2, 11>(__VA_ARGS__);                  \
        case 12: return FN<2, 12>(__VA_ARGS__);                  \
        default: return GF_ERR_UNSUPPORTED;                      \
      }                                                          \
    }                                                            \
  }()
