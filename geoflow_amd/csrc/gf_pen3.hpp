// gf_pen3.hpp -- pencil-per-thread 3-D element operator building block.
//
// A regular element's Laplacian separates by direction,
//     L u = sum_d D_d^T ( M G_dd (D_d u) ),
// so each direction can be handled by threads that own whole pencils along
// it.  An element is worked by N*N threads; thread (a,b) owns
//   * the z-column (i=a, j=b) -- loaded from HBM, z-part done in registers,
//   * the x-pencil (j=a, k=b) and the y-pencil (i=a, k=b) -- read once from
//     the element staged in LDS; gradient, scaling and D^T are all applied
//     in registers.
// D(l,m) entries are wave-uniform in every contraction (the loop indices are
// compile-time) and are compile-time constants (gf_dtab.hpp): the kernels
// run only on meshes whose D is bit-identical to that table, so no thread
// keeps D in VGPRs and no scalar registers are spent holding it.
// LDS traffic is ~7 accesses per node.  The LDS element uses row stride N+1
// and slice stride N(N+1) (= 8 mod 16 doubles for N=8), which makes the
// column writes and both pencil reads (nearly) bank-conflict free.
#pragma once

#include <hip/hip_runtime.h>

#include "gf_dtab.hpp"

namespace gf {

// Six copies of D for the pencil passes: R[d] row-major (R[d][l*N+m] = D(l,m)),
// C[d] column-major (C[d][l*N+m] = D(m,l)), d = x, y, z.
template <typename T>
struct PenD {
  const T *R[3];
  const T *C[3];
};

// Returns p unchanged, but hidden from the optimiser: loads through the
// result cannot be hoisted above this point or merged with earlier loads.
// Used per row of D, so only one row is live in scalar registers at a time
// instead of the whole matrix being hoisted (and spilled) ahead of the loops.
template <typename P>
__device__ __forceinline__ P opaque_ptr(P p) {
  asm volatile("" : "+s"(p));
  return p;
}

// y = A x for the GLL matrix A = D (TR false) or D^T (TR true) at even N,
// through the even/odd split of a centro-antisymmetric matrix
// (A(N-1-l, N-1-m) = -A(l, m), which D and D^T are): with
// e_m = x_m + x_{N-1-m}, o_m = x_m - x_{N-1-m} (m < N/2) and
// P(l, m) = (A(l, m) + A(l, N-1-m)) / 2, Q(l, m) = (A(l, m) - A(l, N-1-m)) / 2,
//   y_l = E_l + O_l,  y_{N-1-l} = O_l - E_l  (l < N/2),
//   E_l = sum_m P(l, m) e_m,  O_l = sum_m Q(l, m) o_m,
// N^2/2 multiply-adds and 2N adds instead of N^2 multiply-adds.  P and Q
// are formed in FP64 from the table (and rounded once to T).  At even N the
// table's D is centro-antisymmetric bit for bit, so y is D x up to the
// rounding of the different operation order (relative 1e-16 in FP64).
template <int N, bool TR, typename T>
__device__ __forceinline__ void gll_apply_eo(const T (&x)[N], T (&y)[N]) {
  static_assert(N % 2 == 0, "even N");
  constexpr int H = N / 2;
  auto A = [](int l, int m) { return TR ? GllD<N>::at(m, l) : GllD<N>::at(l, m); };
  T e[H], o[H];
#pragma unroll
  for (int m = 0; m < H; m++) {
    e[m] = x[m] + x[N - 1 - m];
    o[m] = x[m] - x[N - 1 - m];
  }
#pragma unroll
  for (int l = 0; l < H; l++) {
    T E = T(0), O = T(0);
#pragma unroll
    for (int m = 0; m < H; m++) {
      E = fma((T)(0.5 * (A(l, m) + A(l, N - 1 - m))), e[m], E);
      O = fma((T)(0.5 * (A(l, m) - A(l, N - 1 - m))), o[m], O);
    }
    y[l] = E + O;
    y[N - 1 - l] = O - E;
  }
}

template <int N>
struct Pen3 {
  // Each pass reads D through its own copy (rows contiguous for the gradient,
  // columns contiguous for D^T), so no D value is shared between passes and
  // only one row or column needs to be live in scalar registers at a time.
  static constexpr int NN = N * N * N;
  static constexpr int N2 = N * N;
  static constexpr int P = N + 1;       // row stride (j)
  static constexpr int SK = N * (N + 1);  // slice stride (k)
  static constexpr int SZ = N * SK;       // LDS element size
  static constexpr int EPB = N2 >= 64 ? 1 : 64 / N2;  // elements per workgroup
  static constexpr int THREADS = EPB * N2;
  __device__ static constexpr int at(int i, int j, int k) { return i + j * P + k * SK; }
};

// Computes, for the column (i=a, j=b) owned by the calling thread,
//   post(k, Lu_k, M_k, hx_k, hy_k),   Lu = ((x-part + y-part) + z-part)
// where each part is D_d^T (M * ((D_d u) * G_d) [* p]), the operations of
// GHelmholtz::reg_prod / GGrid::compute_grefdiv (ghelmholtz.ipp:308-378,
// ggrid.ipp:1847-1927), each contraction summed in ascending m at odd N and
// in even/odd form at even N (gll_apply_eo), and M_k is the lumped mass
// W_i W_j W_k Jac (gmass.ipp:200-210; formed here as W_k ((W_j W_i) Jac), which
// rounds differently from the stored mass by at most an ulp or two).
//   ru    the thread's column of u (nodes (a,b,0..N-1))
//   su,sd two LDS element buffers (Pen3<N>::SZ each) owned by this element
//   sx    a third LDS element buffer when HOOK is true, else unused
//   pe    p field restricted to this element, or NULL
//   Wa,Wb the GLL weights W[a], W[b] (loaded once by the caller)
// Hooks see the reference derivatives where they are formed:
//   zcb(k, uz, M)  at column node (a,b,k);
//   xcb(i, ux, M)  at x-pencil node (i,a,b), returns a value hx;
//   ycb(j, uy, M)  at y-pencil node (a,j,b), returns a value hy;
// with HOOK true, hx and hy are carried through LDS to the node's column
// owner and handed to post (advection-type terms); with HOOK false they are
// ignored and post receives zeros.
// All threads of the workgroup must call this (it contains barriers).
template <int N, typename T, bool HOOK, typename XCB, typename YCB, typename ZCB, typename POST>
__device__ __forceinline__ void pen_op3(const T (&ru)[N], T *su, T *sd, T *sx, const PenD<T> &Dp,
                                        const T *__restrict__ W, int a, int b, T Wa, T Wb, T jac,
                                        T g0, T g1,
                                        T g2, const T *__restrict__ pe, XCB &&xcb, YCB &&ycb,
                                        ZCB &&zcb, POST &&post) {
  using S = Pen3<N>;
  (void)Dp;
  // even N: the contractions in even/odd form (gll_apply_eo); odd N: the
  // reference's ascending-m sums
  constexpr bool EO = N % 2 == 0;
  // D(l,m) and D(m,l) as compile-time constants
  auto dR = [](int l, int m) { return (T)GllD<N>::at(l, m); };
  auto dC = [](int l, int m) { return (T)GllD<N>::at(m, l); };
  // lumped mass W_i W_j W_k Jac at the thread's nodes: one product per node
  // (the factor shared by all of them is formed once)
  const T cab = (Wb * Wa) * jac;
#pragma unroll
  for (int k = 0; k < N; k++) su[S::at(a, b, k)] = ru[k];
  // z-part on the column, in registers
  T dz[N];
  {
    T gz[N], uzv[N];
    if constexpr (EO) gll_apply_eo<N, false, T>(ru, uzv);
#pragma unroll
    for (int k = 0; k < N; k++) {
      T uz = T(0);
      if constexpr (EO) {
        uz = uzv[k];
      } else {
#pragma unroll
        for (int m = 0; m < N; m++) uz = fma(dR(k, m), ru[m], uz);
      }
      const T M = W[k] * cab;
      zcb(k, uz, M);
      T f = uz * g2;
      if (pe) f *= pe[a + b * N + k * N * N];
      gz[k] = M * f;
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (EO) {
      gll_apply_eo<N, true, T>(gz, dz);
    } else {
#pragma unroll
      for (int k = 0; k < N; k++) {
        T c = T(0);
#pragma unroll
        for (int m = 0; m < N; m++) c = fma(dC(k, m), gz[m], c);
        dz[k] = c;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  __syncthreads();
  // x-part on the x-pencil (j=a, k=b) -> sd
  {
    T xp[N], f[N], uxv[N];
#pragma unroll
    for (int m = 0; m < N; m++) xp[m] = su[S::at(m, a, b)];
    if constexpr (EO) gll_apply_eo<N, false, T>(xp, uxv);
#pragma unroll
    for (int i = 0; i < N; i++) {
      T ux = T(0);
      if constexpr (EO) {
        ux = uxv[i];
      } else {
#pragma unroll
        for (int m = 0; m < N; m++) ux = fma(dR(i, m), xp[m], ux);
      }
      const T M = W[i] * cab;
      const T hx = xcb(i, ux, M);
      if (HOOK) sx[S::at(i, a, b)] = hx;
      T g = ux * g0;
      if (pe) g *= pe[i + a * N + b * N * N];
      f[i] = M * g;
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (EO) {
      T sv[N];
      gll_apply_eo<N, true, T>(f, sv);
#pragma unroll
      for (int i = 0; i < N; i++) sd[S::at(i, a, b)] = sv[i];
    } else {
#pragma unroll
      for (int i = 0; i < N; i++) {
        T s = T(0);
#pragma unroll
        for (int m = 0; m < N; m++) s = fma(dC(i, m), f[m], s);
        sd[S::at(i, a, b)] = s;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // y-part on the y-pencil (i=a, k=b), kept in registers until sd is complete
  T dy[N], hy[HOOK ? N : 1];
  {
    T yp[N], f[N], uyv[N];
#pragma unroll
    for (int m = 0; m < N; m++) yp[m] = su[S::at(a, m, b)];
    if constexpr (EO) gll_apply_eo<N, false, T>(yp, uyv);
#pragma unroll
    for (int j = 0; j < N; j++) {
      T uy = T(0);
      if constexpr (EO) {
        uy = uyv[j];
      } else {
#pragma unroll
        for (int m = 0; m < N; m++) uy = fma(dR(j, m), yp[m], uy);
      }
      const T M = W[j] * cab;
      const T h = ycb(j, uy, M);
      if (HOOK) hy[j] = h;
      T g = uy * g1;
      if (pe) g *= pe[a + j * N + b * N * N];
      f[j] = M * g;
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (EO) {
      gll_apply_eo<N, true, T>(f, dy);
    } else {
#pragma unroll
      for (int j = 0; j < N; j++) {
        T s = T(0);
#pragma unroll
        for (int m = 0; m < N; m++) s = fma(dC(j, m), f[m], s);
        dy[j] = s;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  __syncthreads();
  // (x-part + y-part) at the y-pencil's nodes; su is free for hy
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int q = S::at(a, j, b);
    sd[q] = sd[q] + dy[j];
    if (HOOK) su[q] = hy[j];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int q = S::at(a, b, k);
    const T M = W[k] * cab;
    post(k, sd[q] + dz[k], M, HOOK ? sx[q] : T(0), HOOK ? su[q] : T(0));
  }
}

struct NoPenCB {
  template <typename T>
  __device__ __forceinline__ T operator()(int, T, T) const {
    return T(0);
  }
};

}  // namespace gf
