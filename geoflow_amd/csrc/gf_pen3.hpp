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
  // 1-D stiffness matrix K = D^T diag(W) D, row-major (FP32 meshes of
  // order 7 only; else null)
  const T *K = nullptr;
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
  // The hook buffer is written along x-pencils and read along columns only.
  // For N = 8 it is unpadded (512 doubles) with the x index XOR-swizzled by
  // (y ^ z): both access patterns then touch 16 distinct 8-byte banks per
  // 16 lanes, and the three LDS buffers of a stage kernel fit 12 waves per
  // CU (3 per SIMD); other orders use the padded layout.
  static constexpr bool HSW = N == 8;
  static constexpr int SZH = HSW ? NN : SZ;
  __device__ static constexpr int hat(int i, int j, int k) {
    return HSW ? (i ^ ((j ^ k) & (N - 1))) + j * N + k * N2 : at(i, j, k);
  }
};

// Computes, for the column (i=a, j=b) owned by the calling thread,
//   post(k, Lu_k, M_k, hx_k, hy_k),   Lu = ((x-part + y-part) + z-part)
// where each part is D_d^T (M * ((D_d u) * G_d) [* p]) summed in ascending m,
// i.e. the order of GHelmholtz::reg_prod / GGrid::compute_grefdiv
// (ghelmholtz.ipp:308-378, ggrid.ipp:1847-1927), and M_k is the lumped mass
// W_i W_j W_k Jac (gmass.ipp:200-210; formed here as W_k ((W_j W_i) Jac), which
// rounds differently from the stored mass by at most an ulp or two).
//   ru    the thread's column of u (nodes (a,b,0..N-1))
//   su,sd two LDS element buffers (Pen3<N>::SZ each) owned by this element
//   sx    a third LDS buffer (Pen3<N>::SZH) when HOOK is true, else unused
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
    T gz[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
      T uz = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) uz = fma(dR(k, m), ru[m], uz);
      const T M = W[k] * cab;
      zcb(k, uz, M);
      T f = uz * g2;
      if (pe) f *= pe[a + b * N + k * N * N];
      gz[k] = M * f;
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      T c = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) c = fma(dC(k, m), gz[m], c);
      dz[k] = c;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();
  // x-part on the x-pencil (j=a, k=b) -> sd
  {
    T xp[N], f[N];
#pragma unroll
    for (int m = 0; m < N; m++) xp[m] = su[S::at(m, a, b)];
#pragma unroll
    for (int i = 0; i < N; i++) {
      T ux = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) ux = fma(dR(i, m), xp[m], ux);
      const T M = W[i] * cab;
      const T hx = xcb(i, ux, M);
      if (HOOK) sx[S::hat(i, a, b)] = hx;
      T g = ux * g0;
      if (pe) g *= pe[i + a * N + b * N * N];
      f[i] = M * g;
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
      T s = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) s = fma(dC(i, m), f[m], s);
      sd[S::at(i, a, b)] = s;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // y-part on the y-pencil (i=a, k=b), kept in registers until sd is complete
  T dy[N], hy[HOOK ? N : 1];
  {
    T yp[N], f[N];
#pragma unroll
    for (int m = 0; m < N; m++) yp[m] = su[S::at(a, m, b)];
#pragma unroll
    for (int j = 0; j < N; j++) {
      T uy = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) uy = fma(dR(j, m), yp[m], uy);
      const T M = W[j] * cab;
      const T h = ycb(j, uy, M);
      if (HOOK) hy[j] = h;
      T g = uy * g1;
      if (pe) g *= pe[a + j * N + b * N * N];
      f[j] = M * g;
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
      T s = T(0);
#pragma unroll
      for (int m = 0; m < N; m++) s = fma(dC(j, m), f[m], s);
      dy[j] = s;
      __builtin_amdgcn_sched_barrier(0);
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
    post(k, sd[q] + dz[k], M, HOOK ? sx[S::hat(a, b, k)] : T(0), HOOK ? su[q] : T(0));
  }
}

// ---------------------------------------------------------------------------
// FP32 Laplacian of a p=7 element (N = 8, one wave per element) in stiffness
// form on the matrix cores.  On a regular element the d-part of L u,
// D_d^T (M G_dd (D_d u)), equals at node (i,j,k), for d = x,
//     (W_j W_k Jac G_xx) sum_l K(i,l) u(l,j,k),   K = D^T diag(W) D,
// the 1-D stiffness matrix (symmetric), and likewise along y and z.  The
// pencils a lane owns -- the z-column (a,b,.), the x-pencil (.,a,b) and
// the y-pencil (a,.,b) -- all have (W_a W_b Jac) as that factor, so each
// direction is ONE batch of 64 length-8 contractions per element (instead
// of a gradient and a D^T batch) followed by one multiply.  Each batch is 16
// rank-1 block updates of v_mfma_f32_4x4x1f32 (16 independent 4x4 blocks
// per wave): block b covers pencils 4b..4b+3 (one per lane of the block;
// the B operand is the lane's own u(l)) and rows 4t..4t+3 of K (the A
// operand K(4t + lane%4, l)), accumulated over l, for t = 0, 1; the result
// lands in the lane that owns the pencil, row 4t + register.  Every block
// product is used and no data moves between lanes.  This regroups the
// reference's sums (ghelmholtz.ipp:308-378), so it serves the FP32 applies
// only, whose results are FP32-accurate either way.
// ---------------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// K in LDS for the whole kernel (64 floats): a lane reads rows 4t + lane%4
// right before the contraction that needs them, so the A operands occupy
// registers only while it runs.
struct MfmaK8 {
  const float *sK;
  int r;  // lane % 4
  // Copies K (global, row-major) into the LDS table; every lane of the
  // wave takes part.  The caller synchronises before the first contraction.
  __device__ __forceinline__ void init(float *lds, const float *__restrict__ K, int lane) {
    lds[lane] = K[lane];
    sK = lds;
    r = lane & 3;
  }
};

// out(l) = sum_m K(l, m) v(m) for the lane's own pencil.  Four independent
// accumulation chains (two row blocks t, two halves of the sum over m),
// interleaved so that consecutive matrix instructions do not wait on each
// other's results; the halves are added at the end.
__device__ __forceinline__ void mfma_contract8(const MfmaK8 &Km, const float (&v)[8],
                                               float (&out)[8]) {
  const float *row0 = Km.sK + Km.r * 8, *row1 = Km.sK + (4 + Km.r) * 8;
  asm volatile("" : "+v"(row0), "+v"(row1));  // keep the reads here
  const float4 a0 = *(const float4 *)row0, a1 = *(const float4 *)(row0 + 4);
  const float4 b0 = *(const float4 *)row1, b1 = *(const float4 *)(row1 + 4);
  const float A0[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const float A1[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  mfma_f4 c00 = {0.f, 0.f, 0.f, 0.f}, c01 = c00, c10 = c00, c11 = c00;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    c00 = __builtin_amdgcn_mfma_f32_4x4x1f32(A0[m], v[m], c00, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_f32_4x4x1f32(A1[m], v[m], c10, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_f32_4x4x1f32(A0[m + 4], v[m + 4], c01, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_f32_4x4x1f32(A1[m + 4], v[m + 4], c11, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    out[q] = c00[q] + c01[q];
    out[4 + q] = c10[q] + c11[q];
  }
}

// post(k, Lu_k, M_k, 0, 0) for the lane's column, as in pen_op3, with M_k =
// W_k ((W_b W_a) Jac).  One element per wave: lane = a + 8 b.  All lanes of
// the workgroup must call this (barriers).
template <typename POST>
__device__ __forceinline__ void pen_lap3k_mfma8(const float (&ru)[8], float *su, float *sd,
                                                const MfmaK8 &Km, const float *__restrict__ W,
                                                float Wa, float Wb, int a, int b, float jac,
                                                float g0, float g1, float g2, POST &&post) {
  using S = Pen3<8>;
  const float cab = (Wb * Wa) * jac;
#pragma unroll
  for (int k = 0; k < 8; k++) su[S::at(a, b, k)] = ru[k];
  float kz[8];
  mfma_contract8(Km, ru, kz);
  __syncthreads();
  {
    float xp[8], kx[8];
#pragma unroll
    for (int m = 0; m < 8; m++) xp[m] = su[S::at(m, a, b)];
    mfma_contract8(Km, xp, kx);
    const float fx = cab * g0;
#pragma unroll
    for (int i = 0; i < 8; i++) sd[S::at(i, a, b)] = fx * kx[i];
  }
  float ky[8];
  {
    float yp[8];
#pragma unroll
    for (int m = 0; m < 8; m++) yp[m] = su[S::at(a, m, b)];
    mfma_contract8(Km, yp, ky);
  }
  __syncthreads();
  const float fy = cab * g1;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int q = S::at(a, j, b);
    sd[q] = sd[q] + fy * ky[j];
  }
  __syncthreads();
  const float fz = cab * g2;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int q = S::at(a, b, k);
    post(k, sd[q] + fz * kz[k], W[k] * cab, 0.f, 0.f);
  }
}

struct NoPenCB {
  template <typename T>
  __device__ __forceinline__ T operator()(int, T, T) const {
    return T(0);
  }
};

}  // namespace gf
