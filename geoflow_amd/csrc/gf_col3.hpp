// gf_col3.hpp -- column-per-thread 3-D element operator building block.
//
// One element is handled by N*N threads; thread (i,j) owns the z-column of
// nodes (i,j,0..N-1) in registers.  Per z-slice k the slice of the operand is
// staged in LDS (row stride N+1 so the 8-byte rows of a wave64 fall in
// distinct banks), the x and y contractions read it, and the z contraction
// runs entirely in registers with wave-uniform D(k,m) operands.  Each thread
// keeps the rows D(i,:), D(j,:) and columns D(:,i), D(:,j) it needs in VGPRs.
// LDS holds three slices per element (operand, x-flux, y-flux).
#pragma once

#include <hip/hip_runtime.h>

namespace gf {

template <int N>
struct Col3 {
  static constexpr int NN = N * N * N;
  static constexpr int N2 = N * N;
  static constexpr int P = N + 1;              // padded LDS row stride
  static constexpr int SL = N * P;             // one padded slice
  static constexpr int EPB = N2 >= 64 ? 1 : 64 / N2;  // elements per workgroup
  static constexpr int THREADS = EPB * N2;
};

// Register budget: 4 waves per SIMD (<= 128 VGPRs) so that LDS and HBM
// latencies are hidden.
template <typename T>
struct ColWaves {
  static constexpr int value = 4;
};

// Per-thread copies of the derivative-matrix rows D(i,:) and D(j,:).
template <int N, typename T>
struct DRegs {
  T xr[N], yr[N];
  __device__ __forceinline__ void load(const T *__restrict__ D, int i, int j) {
#pragma unroll
    for (int m = 0; m < N; m++) {
      xr[m] = D[i + m * N];  // D(i,m)
      yr[m] = D[j + m * N];  // D(j,m)
    }
  }
};

// LDS copies of D shared by a workgroup: sDt[l*(N+1) + m] = D(m,l) (columns
// as padded rows, read per lane without bank conflicts) and sDz[l*N + m] =
// D(l,m) (rows, read wave-uniformly as broadcasts).  Call stage() by every
// thread, followed by a barrier before use.
template <int N, typename T>
struct DShared {
  T t[N * (N + 1)];
  T z[N * N];
  __device__ __forceinline__ void stage(const T *__restrict__ D) {
    for (int q = threadIdx.x; q < N * N; q += blockDim.x) {
      int l = q % N, m = q / N;  // D(l,m) = D[l + m*N]
      T v = D[q];
      t[m * (N + 1) + l] = v;    // row m of t holds column m of D
      z[l * N + m] = v;
    }
  }
};

// Element operator  L u = sum_d D_d^T ( M * ((D_d u) * G_d) [* p] )  at the
// thread's column: the reference-space divergence of the scaled gradient,
// formed as ((x-part + y-part) + z-part) with each contraction summed in
// ascending m -- the order of GHelmholtz::reg_prod / GGrid::compute_grefdiv
// (ghelmholtz.ipp:308-378, ggrid.ipp:1847-1927).
//
// Pass 1 forms the z-fluxes in registers; pass 2 walks the z-slices, stages
// each slice of u in LDS for the x/y contractions and hands every finished
// node to `post`, so no per-node output array stays live.  Hooks:
//   zcb(k, uz, M)            reference z-derivative and lumped mass (pass 1)
//   xycb(k, ux, uy, M)       reference x/y-derivatives (pass 2)
//   post(k, Lu, M)           operator value at node (i,j,k) (pass 2)
// All threads of the workgroup must call this (it contains barriers).
template <int N, typename T, typename ZCB, typename XYCB, typename POST>
__device__ __forceinline__ void col_op3(const T (&ru)[N], T *su, T *sgx, T *sgy,
                                        const DRegs<N, T> &dr, const DShared<N, T> &ds,
                                        const T *__restrict__ W, int i, int j, T jac, T g0, T g1,
                                        T g2, const T *__restrict__ pcol, ZCB &&zcb, XYCB &&xycb,
                                        POST &&post) {
  constexpr int P = N + 1;
  const T Wi = W[i], Wj = W[j];
  T gz[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    T uz = T(0);
#pragma unroll
    for (int m = 0; m < N; m++) uz = fma(ds.z[k * N + m], ru[m], uz);
    const T M = W[k] * Wj * Wi * jac;  // ((W_k W_j) W_i) Jac, gmass.ipp:200-210
    zcb(k, uz, M);
    T fz = uz * g2;
    if (pcol) fz *= pcol[k * N * N];
    gz[k] = M * fz;
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    su[j * P + i] = ru[k];
    __syncthreads();
    T ux = T(0), uy = T(0);
#pragma unroll
    for (int m = 0; m < N; m++) ux = fma(dr.xr[m], su[j * P + m], ux);
#pragma unroll
    for (int m = 0; m < N; m++) uy = fma(dr.yr[m], su[m * P + i], uy);
    const T M = W[k] * Wj * Wi * jac;
    xycb(k, ux, uy, M);
    T fx = ux * g0, fy = uy * g1;
    if (pcol) {
      const T pv = pcol[k * N * N];
      fx *= pv;
      fy *= pv;
    }
    sgx[j * P + i] = M * fx;
    sgy[j * P + i] = M * fy;
    __syncthreads();
    T a = T(0), b = T(0), c = T(0);
#pragma unroll
    for (int m = 0; m < N; m++) a = fma(ds.t[i * P + m], sgx[j * P + m], a);
#pragma unroll
    for (int m = 0; m < N; m++) b = fma(ds.t[j * P + m], sgy[m * P + i], b);
#pragma unroll
    for (int m = 0; m < N; m++) c = fma(ds.z[m * N + k], gz[m], c);
    post(k, (a + b) + c, M);
  }
}

struct NoZ {
  template <typename T>
  __device__ __forceinline__ void operator()(int, T, T) const {}
};
struct NoXY {
  template <typename T>
  __device__ __forceinline__ void operator()(int, T, T, T) const {}
};


}  // namespace gf
