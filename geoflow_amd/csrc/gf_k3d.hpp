// gf_k3d.hpp -- 3-D regular-element kernels (pencil form, gf_pen3.hpp) and
// their launchers: Helmholtz apply, CG apply and update, Burgers RK stages
// and the structured-box RK output.  Included by the gf_k3d_n*.hip
// translation units, each of which instantiates the launchers for some
// element orders (GF_K3D_INSTANTIATE).
#pragma once

#include <type_traits>

#include "gf_solve_common.hpp"

namespace gfs {

// ---------------------------------------------------------------------------
// Pencil-form 3-D kernels (see gf_pen3.hpp): Helmholtz apply and CG apply.
// ---------------------------------------------------------------------------
template <int N, typename T>
__global__ __launch_bounds__(gf::Pen3<N>::THREADS) void k_helm3p(
    const T *__restrict__ u, T *__restrict__ uo, gf::PenD<T> D, const T *__restrict__ W,
    const T *__restrict__ edx, const T *__restrict__ ejac, T p, const T *__restrict__ pf,
    const T *__restrict__ qf, T qs, size_t nelems) {
  using S = gf::Pen3<N>;
  __shared__ T su[S::EPB][S::SZ], sd[S::EPB][S::SZ];
  ColIdx<N> t(nelems);
  const size_t col = t.e * S::NN + t.r;
  T ru[N];
#pragma unroll
  for (int k = 0; k < N; k++) ru[k] = u[col + k * S::N2];
  const T jac = ejac[t.e];
  const T f0 = edx[t.e * 3], f1 = edx[t.e * 3 + 1], f2 = edx[t.e * 3 + 2];
  auto post = [&](int k, T Lu, T M, T, T) {
    if (!t.active) return;
    T acc = Lu;
    if (!pf && p != T(1)) acc *= p;
    if (qf) {
      T tq = M * ru[k];
      tq *= qf[col + k * S::N2];
      acc += tq;
    } else if (qs != T(0)) {
      T tq = M * ru[k];
      if (qs != T(1)) tq *= qs;
      acc += tq;
    }
    uo[col + k * S::N2] = acc;
  };
  gf::pen_op3<N, T, false>(ru, su[t.el], sd[t.el], (T *)nullptr, D, W, t.i, t.j, W[t.i], W[t.j],
                           jac, f0 * f0,
                           f1 * f1, f2 * f2, pf ? pf + t.e * S::NN : (const T *)nullptr,
                           gf::NoPenCB(), gf::NoPenCB(), gf::NoPenCB(), post);
}

template <int N, typename T>
__global__ __launch_bounds__(gf::Pen3<N>::THREADS) void k_cg_apply3p(
    const T *__restrict__ r, const T *__restrict__ dinv, T *__restrict__ w, T *__restrict__ qv,
    gf::PenD<T> Dp, const T *__restrict__ W, const T *__restrict__ edx,
    const T *__restrict__ ejac, const T *__restrict__ qf, T qs, const T *__restrict__ mask, double pd,
    const double *__restrict__ scal, double *__restrict__ part, size_t nelems, gf::BoxTopo topo,
    int tmask, const int64_t *__restrict__ seg_eoff, int nseg, const double *__restrict__ sscal,
    const T *__restrict__ dtab, const unsigned char *__restrict__ deq) {
  using S = gf::Pen3<N>;
  __shared__ T su[S::EPB][S::SZ], sd[S::EPB][S::SZ];
  __shared__ double sred[S::THREADS > 64 ? S::THREADS : 64];
  ColIdx<N> t(nelems);
  // batched mesh: each element's member scalars; an element of a member
  // that has converged computes nothing it stores, and the partial sums are
  // kept per element so that each member's are one contiguous range
  if (nseg > 0) {
    scal = sscal + kSegScal * seg_of(seg_eoff, nseg, t.e);
    if (scal[S_DONE] != 0.0) {
      if (S::EPB == 1) return;  // the whole workgroup
      t.active = false;
    }
  } else if (scal[S_DONE] != 0.0) {
    return;
  }
  const bool first = scal[S_FIRST] != 0.0;
  const T beta = (T)scal[S_BETA];
  const size_t col = t.e * S::NN + t.r;
  const T jac = ejac[t.e];
  const T f0 = edx[t.e * 3], f1 = edx[t.e * 3 + 1], f2 = edx[t.e * 3 + 2];
  // all loads of the column first, then the stores of w; the Jacobi values
  // from the table where the element's equal it
  const T *dsrc = dinv ? ((dtab && deq[t.e]) ? dtab + t.r : dinv + col) : nullptr;
  T ru[N], dv[N], wv[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    const size_t g = col + k * S::N2;
    ru[k] = r[g];
    dv[k] = dsrc ? dsrc[k * S::N2] : T(1);
    wv[k] = first ? T(0) : w[g];
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    const T z = dinv ? dv[k] * ru[k] : ru[k];
    ru[k] = first ? z : beta * wv[k] + T(1) * z;
  }
  if (t.active) {
#pragma unroll
    for (int k = 0; k < N; k++) w[col + k * S::N2] = ru[k];
  }
  double dot = 0.0;
  const T p = (T)pd;
  int ec[3] = {0, 0, 0};
  if (tmask) gf::box_elem_coords(topo, t.e, ec);
  auto post = [&](int k, T Lu, T M, T, T) {
    if (!t.active) return;
    const size_t g = col + k * S::N2;
    T acc = Lu;
    if (p != T(1)) acc *= p;
    if (qf) {
      T tq = M * ru[k];
      tq *= qf[g];
      acc += tq;
    } else if (qs != T(0)) {
      T tq = M * ru[k];
      if (qs != T(1)) tq *= qs;
      acc += tq;
    }
    if (mask) acc *= mask[g];
    else if (tmask && gf::box_masked(topo, N, ec, t.i, t.j, k)) acc *= T(0);
    qv[g] = acc;
    dot += (double)ru[k] * (double)acc;
  };
  gf::pen_op3<N, T, false>(ru, su[t.el], sd[t.el], (T *)nullptr, Dp, W, t.i, t.j, W[t.i],
                           W[t.j], jac, f0 * f0, f1 * f1, f2 * f2, (const T *)nullptr,
                           gf::NoPenCB(), gf::NoPenCB(), gf::NoPenCB(), post);
  if (nseg > 0) {
    // per-element partials: the element's lanes, summed in lane order
    sred[threadIdx.x] = dot;
    __syncthreads();
    if (t.r == 0 && t.active) {
      double v = 0.0;
      for (int q2 = 0; q2 < S::N2; q2++) v += sred[t.el * S::N2 + q2];
      part[t.e] = v;
    }
    return;
  }
  double bs = gf::block_reduce<false>(dot, sred);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// Fused Burgers stage, 3-D pencil form (see k_burgers_stage for the contract).
// The x- and y-derivative factors of the advection term are carried to the
// column owner through LDS, so every node's advection is summed in the
// reference order ((ux rx) u0 + (uy ry) u1) + (uz rz) u2 (gadvect.ipp:58-100).
template <int N>
__global__ __launch_bounds__(gf::Pen3<N>::THREADS) __attribute__((amdgpu_waves_per_eu(N <= 8 ? 3 : 1)))
void k_burgers_stage3p(
    StageArgs a, gf::PenD<double> Dp) {
  using S = gf::Pen3<N>;
  __shared__ double su[S::EPB][S::SZ], sd[S::EPB][S::SZ], sx[S::EPB][S::SZ];
  ColIdx<N> t(a.nelems);
  const size_t col = t.e * S::NN + t.r;
  // the inverse mass of the column, read once for all components
  double im[N];
#pragma unroll
  for (int k = 0; k < N; k++) im[k] = a.imass[col + k * S::N2];
  double ru[3][N];
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      const size_t g = col + k * S::N2;
      const double ui = a.uin[c][g];
      double v = ui;
      if (a.kprev[c]) {
        const double kp = a.kprev[c][g];
        v = ui + a.bh * kp;
        if (a.acc[c] && t.active) a.acc[c][g] = a.accsrc[c][g] + kp * a.ch;
      }
      ru[c][k] = v;
    }
  }
  const double jac = a.ejac[t.e];
  const double wa = a.W[t.i], wb = a.W[t.j];
  const double rx = a.edx[t.e * 3], ry = a.edx[t.e * 3 + 1], rz = a.edx[t.e * 3 + 2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    // the z pass's derivatives, kept for the advection term
    double uzs[N];
    auto zcb = [&](int k, double uz, double) { uzs[k] = uz; };
    auto xcb = [&](int, double ux, double) { return ux * rx; };
    auto ycb = [&](int, double uy, double) { return uy * ry; };
    double *kout = a.kout[c];
    auto post = [&](int k, double L, double M, double hx, double hy) {
      if (!t.active) return;
      const size_t g = col + k * S::N2;
      const double uz = uzs[k];
      double po = hx * ru[0][k];
      po += hy * ru[1][k];
      po += (uz * rz) * ru[2][k];
      const double adv = po * M;
      if (a.nu != 1.0) L *= a.nu;
      const double urhs = -1.0 * L + -1.0 * adv;
      kout[g] = im[k] * urhs;
    };
    gf::pen_op3<N, double, true>(ru[c], su[t.el], sd[t.el], sx[t.el], Dp, a.W, t.i, t.j, wa, wb, jac,
                                 rx * rx, ry * ry, rz * rz, (const double *)nullptr, xcb, ycb, zcb,
                                 post);
    __syncthreads();  // the next component reuses the LDS buffers
  }
}

template <int N>
int launch_stage3p(gf_mesh *m, const StageArgs &a) {
  using S = gf::Pen3<N>;
  int nb = xcd_grid((m->nelems + S::EPB - 1) / S::EPB);
  hipLaunchKernelGGL((k_burgers_stage3p<N>), dim3(nb), dim3(S::THREADS), 0, m->ctx->stream, a,
                     pend_d(m));
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// ---------------------------------------------------------------------------
// Burgers RK on a structured box (mesh->has_box): the GGFX Smooth of each
// stage derivative, of the final-stage input and of the step output is formed
// by gather in the kernels that consume them (gf_boxdss.hpp), in the GGFX
// summation order, instead of separate DSS passes.  Stage derivatives are
// stored unsmoothed.
//   kp       previous stage derivative K_{s-1} (unsmoothed) or NULL (stage 1)
//   acc_mode 0: no accumulator; 1: acc = Smooth(uin) + Ks*ch; 2: acc += Ks*ch
//            (step_b accumulates from uin and smooths the output at the end;
//            starting from Smooth(uin) keeps the accumulator continuous, so
//            the final Smooth is the identity up to rounding and the output
//            kernel needs only the last stage derivative's group sums)
//   smooth_v the stage input is the final-stage input of step_b, which is
//            smoothed: v = Smooth(uin + bh*Ks) (gexrk_stepper.ipp:300-318)
//   acc_mode 3: the previous step's output fused with this step's first
//            stage: uin is the previous accumulator, bh the last weight c h,
//            and v = acc + c h Smooth(K_last) -- the output kernel's value,
//            in the same operations -- is stored through `acc` (the state)
//            and is this stage's input
// ---------------------------------------------------------------------------
template <int N, bool KP, int ACC, bool SMV, bool HL>
__global__ __launch_bounds__(gf::Pen3<N>::THREADS) __attribute__((amdgpu_waves_per_eu(N == 8 ? 4 : 1)))
void k_burgers_stage3b(const double *__restrict__ ui0, const double *__restrict__ ui1,
                       const double *__restrict__ ui2, const double *__restrict__ kp0,
                       const double *__restrict__ kp1, const double *__restrict__ kp2,
                       double *__restrict__ ac0, double *__restrict__ ac1, double *__restrict__ ac2,
                       double *__restrict__ ko0, double *__restrict__ ko1, double *__restrict__ ko2,
                       const double *__restrict__ imass, const double *__restrict__ W,
                       const double *__restrict__ edx, const double *__restrict__ ejac,
                       gf::PenD<double> Dp, double bh, double ch, double nu, gf::BoxTopo topo,
                       size_t nelems, size_t imstride) {
  using S = gf::Pen3<N>;
  // N = 8 with a previous stage: the three LDS element buffers share memory
  // with the halos of the gather (lds_copy512, gf::Halo8Rim; the kernel is
  // held to two waves per SIMD by its registers, so 16 KB of LDS per wave
  // costs nothing)
  constexpr bool HALO = N == 8 && KP && HL;
  constexpr bool GU = SMV || ACC == 1;  // the stage reads uin's group sums
  // halo path: K's halo, uin's halo or own values, and (without uin's halo)
  // the accumulator's own values, one after the other
  constexpr int NHL = GU ? 2 * gf::kHalo8 : gf::kHalo8 + 512 + (ACC == 2 ? 512 : 0);
  constexpr int NSL = HALO && NHL > 3 * S::SZ ? NHL : 3 * S::EPB * S::SZ;
  static_assert(!HALO || 3 * S::SZ <= NSL, "element buffers fit the halo area");
  __shared__ double sl[NSL];
  double(*su)[S::SZ] = (double(*)[S::SZ])sl;
  double(*sd)[S::SZ] = (double(*)[S::SZ])(sl + S::EPB * S::SZ);
  double(*sx)[S::SZ] = (double(*)[S::SZ])(sl + 2 * S::EPB * S::SZ);
  const double *uin[3] = {ui0, ui1, ui2};
  const double *kp[3] = {kp0, kp1, kp2};
  double *acc[3] = {ac0, ac1, ac2};
  double *kout[3] = {ko0, ko1, ko2};
  ColIdx<N> t(nelems);
  const size_t col = t.e * S::NN + t.r;
  gf::BoxElemDeltas de = {{0, 0, 0}, {0, 0, 0}};
  if (KP) {
    int ec[3];
    gf::box_elem_coords(topo, t.e, ec);
    de = gf::box_elem_deltas(topo, N, ec);
  }
  // The inverse mass of the column, read once for all components (through
  // LDS on the halo path).
  double im[N];
  if constexpr (HALO) {
    lds_copy512(imass + t.e * imstride, sl, (int)threadIdx.x);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) im[k] = sl[threadIdx.x + 64 * k];
  } else {
#pragma unroll
    for (int k = 0; k < N; k++) im[k] = imass[t.e * imstride + t.r + (size_t)k * S::N2];
  }
  double ru[3][N];
  if constexpr (HALO) {
    // Per component: every load of the previous stage derivative (and of
    // uin's halo where its group sums are needed) is issued at once -- the
    // element's own values copied straight to LDS, the partner values
    // through registers -- and each node sums its group from LDS in GGFX
    // order (halo8_order): the operations and order of the gather path
    // below, so the same result bit for bit.
    const int lane = (int)threadIdx.x;
    const int64_t ebase = (int64_t)t.e * 512;
    // K's halo; uin's halo (GU) or own values; the accumulator's own values
    double *hk = sl, *hu = sl + gf::kHalo8, *ha = sl + gf::kHalo8 + 512;
    const gf::Halo8Col hc(de, t.i, t.j);
    gf::static_for<0, 3>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      __syncthreads();  // the previous component's group sums have read LDS
      // the element's own values straight to LDS, the partner values
      // through registers
      lds_copy512(kp[c] + ebase, hk, lane);
      lds_copy512(uin[c] + ebase, hu, lane);
      if (ACC == 2 && !GU) lds_copy512(acc[c] + ebase, ha, lane);
      gf::Halo8Rim<double> Lk, Lu;
      Lk.issue(kp[c], de, ebase, lane);
      if (GU) Lu.issue(uin[c], de, ebase, lane);
      double ao[8];
#pragma unroll
      for (int k = 0; k < 8; k++) ao[k] = ACC == 2 && GU ? acc[c][col + (size_t)k * 64] : 0.0;
      Lk.store(hk, lane);
      if (GU) Lu.store(hu, lane);
      __syncthreads();
      gf::static_for<0, 8>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int NT = (k == 0 || k == 7) ? 8 : 4;  // terms that can be non-zero
        const int own = lane + 64 * k;
        int ad[8];
        const int cnt = gf::halo8_order<k>(hc, de, ad);
        double sk = 0.0;
#pragma unroll
        for (int m = 0; m < NT; m++) sk = sk + hk[ad[m]];
        // GGFX Smooth: Sum / count (an exact multiply for counts 1, 2, 4, 8)
        const double ks = cnt == 1 ? hk[own] : sk * gf::box_inv(cnt);
        const double uo = hu[own];
        double v = uo + bh * ks;
        double av = 0.0;
        if (ACC == 2) av = (GU ? ao[k] : ha[own]) + ks * ch;
        if (ACC == 3) av = v;
        if (ACC == 1) {
          // the accumulator starts from Smooth(uin), so it stays continuous
          double su1 = 0.0;
#pragma unroll
          for (int m = 0; m < NT; m++) su1 = su1 + hu[ad[m]];
          const double us = cnt == 1 ? uo : su1 * gf::box_inv(cnt);
          av = us + ks * ch;
        }
        if (SMV && cnt > 1) {
          // the members only: an absent partner's zero slot would add bh ks
          double sv = 0.0;
#pragma unroll
          for (int m = 0; m < NT; m++)
            if (ad[m] != gf::kHalo8Zero) sv = sv + (hu[ad[m]] + bh * ks);
          v = sv * gf::box_inv(cnt);
        }
        if (ACC && t.active) acc[c][col + (size_t)k * 64] = av;
        ru[c][k] = v;
      });
    });
    __syncthreads();  // the element buffers below reuse the halo area
  } else {
  // Stage input (and accumulator) per node; all loads are issued before any
  // store, the accumulator stores follow the loop.  Absent group members are
  // not read (box_load8m).
  double av[ACC ? 3 : 1][ACC ? N : 1];
#pragma unroll
  for (int k = 0; k < N; k++) {
    const size_t g = col + (size_t)k * S::N2;
    if (!KP) {
#pragma unroll
      for (int c = 0; c < 3; c++) ru[c][k] = uin[c][g];
      continue;
    }
    int64_t mem[8];
    int own;
    const int cnt = gf::box_members_own<int64_t>(de, N, (int64_t)g, t.i, t.j, k, mem, &own);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double kv[8];
      gf::box_load8m<int64_t, double>(kp[c], mem, cnt, kv);
      const double ao = ACC == 2 ? acc[c][g] : 0.0;
      double uv[8];
      double ui;
      if (GU) gf::box_load8m<int64_t, double>(uin[c], mem, cnt, uv);
      if (SMV) {
        // the node's own value is one of the gathered slots (picking it
        // there measured faster than a separate load in this form only)
        ui = uv[0];
#pragma unroll
        for (int m = 1; m < 8; m++) ui = own == m ? uv[m] : ui;
      } else {
        ui = uin[c][g];
      }
      // GGFX Smooth: Sum / count (an exact multiply for counts 1, 2, 4, 8)
      const double ks = cnt == 1 ? kv[0] : gf::box_sum8z<double>(kv, cnt) * gf::box_inv(cnt);
      double v = ui + bh * ks;
      if (ACC == 2) av[c][k] = ao + ks * ch;
      if (ACC == 3) av[c][k] = v;
      if (ACC == 1) {
        // the accumulator starts from Smooth(uin), so it stays continuous
        const double us = cnt == 1 ? uv[0] : gf::box_sum8z<double>(uv, cnt) * gf::box_inv(cnt);
        av[c][k] = us + ks * ch;
      }
      if (SMV) {
#pragma unroll
        for (int q = 0; q < 8; q++) uv[q] = uv[q] + bh * ks;
        if (cnt > 1) v = gf::box_sum8<double>(uv, cnt) * gf::box_inv(cnt);
      }
      ru[c][k] = v;
    }
  }
  if (ACC && t.active) {
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int k = 0; k < N; k++) acc[c][col + (size_t)k * S::N2] = av[ACC ? c : 0][ACC ? k : 0];
  }
  }
  const double jac = ejac[t.e];
  const double wa = W[t.i], wb = W[t.j];
  const double rx = edx[t.e * 3], ry = edx[t.e * 3 + 1], rz = edx[t.e * 3 + 2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    // the z pass's derivatives, kept for the advection term
    double uzs[N];
    auto zcb = [&](int k, double uz, double) { uzs[k] = uz; };
    auto xcb = [&](int, double ux, double) { return ux * rx; };
    auto ycb = [&](int, double uy, double) { return uy * ry; };
    double *ko = kout[c];
    auto post = [&](int k, double L, double M, double hx, double hy) {
      if (!t.active) return;
      const size_t g = col + k * S::N2;
      const double uz = uzs[k];
      double po = hx * ru[0][k];
      po += hy * ru[1][k];
      po += (uz * rz) * ru[2][k];
      const double adv = po * M;
      if (nu != 1.0) L *= nu;
      const double urhs = -1.0 * L + -1.0 * adv;
      ko[g] = im[k] * urhs;
    };
    gf::pen_op3<N, double, true>(ru[c], su[t.el], sd[t.el], sx[t.el], Dp, W, t.i, t.j, wa, wb, jac,
                                 rx * rx, ry * ry, rz * rz, (const double *)nullptr, xcb, ycb, zcb,
                                 post);
    __syncthreads();  // the next component reuses the LDS buffers
  }
}

// Step output on a structured box: u = acc + cl*Smooth(K), the smooth by
// gather.  step_b forms Smooth(acc + cl*Smooth(K)) (its last update and GGFX
// Smooth); the accumulator here is continuous (it starts from Smooth(uin)
// and adds smoothed stage derivatives), so that final Smooth is the identity
// up to rounding and is not formed.
template <int N, typename I>
__global__ __launch_bounds__(kThreads) void k_rk_out_box(
    double *__restrict__ u0, double *__restrict__ u1, double *__restrict__ u2,
    const double *__restrict__ a0, const double *__restrict__ a1, const double *__restrict__ a2,
    const double *__restrict__ k0, const double *__restrict__ k1, const double *__restrict__ k2,
    double cl, gf::BoxTopo topo, size_t nelems) {
  constexpr int NN = N * N * N;
  constexpr int CE = NN >= kThreads ? 1 : kThreads / NN;
  constexpr int CN = CE * NN;
  double *u[3] = {u0, u1, u2};
  const double *acc[3] = {a0, a1, a2};
  const double *kf[3] = {k0, k1, k2};
  const size_t nchunks = (nelems + CE - 1) / CE;
  const size_t cpb = (nchunks + gridDim.x - 1) / gridDim.x;
  const size_t c0 = xcd_block() * cpb, c1 = c0 + cpb < nchunks ? c0 + cpb : nchunks;
  int ecw[3] = {0, 0, 0};
  if (CE == 1) gf::box_elem_coords(topo, c0, ecw);
  for (size_t ch = c0; ch < c1; ch++) {
    int ec0[3];
    if (CE == 1) {
      ec0[0] = ecw[0];
      ec0[1] = ecw[1];
      ec0[2] = ecw[2];
      gf::box_elem_next(topo, ecw);
    } else {
      gf::box_elem_coords(topo, ch * CE, ec0);
    }
    const gf::BoxElemDeltas d0 = gf::box_elem_deltas(topo, N, ec0);
#pragma unroll
    for (int q0 = 0; q0 < CN; q0 += kThreads) {
      const int q = q0 + (int)threadIdx.x;
      if (q >= CN) break;
      const int el = CE == 1 ? 0 : q / NN;
      const int rr = CE == 1 ? q : q - el * NN;
      const size_t e = ch * CE + el;
      if (e >= nelems) break;
      gf::BoxElemDeltas de = d0;
      if (CE > 1 && el) {
        int ec[3];
        gf::box_elem_coords(topo, e, ec);
        de = gf::box_elem_deltas(topo, N, ec);
      }
      const int i = rr % N, j = (rr / N) % N, k = rr / (N * N);
      const I g = (I)(e * NN + rr);
      I mem[8];
      const int cnt = gf::box_members<I>(de, N, g, i, j, k, mem);
      // every load of the node (all components) before the first store;
      // absent group members are not read
      double kv[3][8], av[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        gf::box_load8m<I, double>(kf[c], mem, cnt, kv[c]);
        av[c] = acc[c][g];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double ks =
            cnt == 1 ? kv[c][0] : gf::box_sum8z<double>(kv[c], cnt) * gf::box_inv(cnt);
        u[c][g] = 1.0 * av[c] + cl * ks;
      }
    }
  }
}

// Step output on an N = 8 box through the LDS halo: one element per
// 64-thread block; per component the element's K and accumulator are copied
// to LDS (lds_copy512) while its partner values are loaded into
// registers (gf::Halo8Rim), and each node sums its group from LDS in canonical order
// (gf::halo8_order) -- k_rk_out_box's operations, the same result bit for
// bit, with every K value fetched once instead of once per group copy.
template <int N>
__global__ __launch_bounds__(64) void k_rk_out_halo8(
    double *__restrict__ u0, double *__restrict__ u1, double *__restrict__ u2,
    const double *__restrict__ a0, const double *__restrict__ a1, const double *__restrict__ a2,
    const double *__restrict__ k0, const double *__restrict__ k1, const double *__restrict__ k2,
    double cl, gf::BoxTopo topo, size_t nelems) {
  // K's halo, then the accumulator's own values
  __shared__ double sl[gf::kHalo8 + 512];
  double *hk = sl, *ha = sl + gf::kHalo8;
  double *u[3] = {u0, u1, u2};
  const double *acc[3] = {a0, a1, a2};
  const double *kf[3] = {k0, k1, k2};
  const size_t e = xcd_block();
  if (e >= nelems) return;  // the whole block
  const int lane = (int)threadIdx.x;
  int ec[3];
  gf::box_elem_coords(topo, e, ec);
  static_assert(N == 8, "the LDS halo serves N = 8");
  const gf::BoxElemDeltas de = gf::box_elem_deltas(topo, 8, ec);
  const gf::Halo8Col hc(de, lane & 7, lane >> 3);
  const int64_t ebase = (int64_t)e * 512;
  gf::static_for<0, 3>([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    __syncthreads();  // the previous component's sums have read LDS
    // own values straight to LDS, the partner values through registers
    lds_copy512(kf[c] + ebase, hk, lane);
    lds_copy512(acc[c] + ebase, ha, lane);
    gf::Halo8Rim<double> L;
    L.issue(kf[c], de, ebase, lane);
    L.store(hk, lane);
    __syncthreads();
    gf::static_for<0, 8>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      constexpr int NT = (k == 0 || k == 7) ? 8 : 4;  // terms that can be non-zero
      const int own = lane + 64 * k;
      int ad[8];
      const int cnt = gf::halo8_order<k>(hc, de, ad);
      double sk = 0.0;
#pragma unroll
      for (int m = 0; m < NT; m++) sk = sk + hk[ad[m]];
      const double ks = cnt == 1 ? hk[own] : sk * gf::box_inv(cnt);
      u[c][ebase + own] = 1.0 * ha[own] + cl * ks;
    });
  });
}

template <int N, bool KP, int ACC, bool SMV, bool HL>
void launch_stage3b_h(gf_mesh *m, int nb, double *const *uin, double *const *kp,
                      double *const *acc, double *const *kout, double bh, double ch, double nu) {
  using S = gf::Pen3<N>;
  hipLaunchKernelGGL((k_burgers_stage3b<N, KP, ACC, SMV, HL>), dim3(nb), dim3(S::THREADS), 0,
                     m->ctx->stream, uin[0], uin[1], uin[2], kp ? kp[0] : nullptr,
                     kp ? kp[1] : nullptr, kp ? kp[2] : nullptr, acc ? acc[0] : nullptr,
                     acc ? acc[1] : nullptr, acc ? acc[2] : nullptr, kout[0], kout[1], kout[2],
                     m->imass_el ? m->imass_el : m->imass, m->W[0], m->edx, m->ejac, pend_d(m),
                     bh, ch, nu, m->box, m->nelems, m->imass_el ? (size_t)0 : (size_t)S::NN);
}

// The N = 8 gather through the LDS halo where the box's partner offsets
// order x < y < z (gf_mesh::halo8), else the per-node gather.
template <int N, bool KP, int ACC, bool SMV>
void launch_stage3b_v(gf_mesh *m, int nb, double *const *uin, double *const *kp,
                      double *const *acc, double *const *kout, double bh, double ch, double nu) {
  if constexpr (N == 8 && KP) {
    if (m->halo8) {
      launch_stage3b_h<N, KP, ACC, SMV, true>(m, nb, uin, kp, acc, kout, bh, ch, nu);
      return;
    }
  }
  launch_stage3b_h<N, KP, ACC, SMV, false>(m, nb, uin, kp, acc, kout, bh, ch, nu);
}

template <int N>
int launch_stage3b(gf_mesh *m, double *const *uin, double *const *kp, double *const *acc,
                   double *const *kout, double bh, double ch, double nu, int acc_mode,
                   int smooth_v) {
  using S = gf::Pen3<N>;
  int nb = xcd_grid((m->nelems + S::EPB - 1) / S::EPB);
  if (!kp)
    launch_stage3b_v<N, false, 0, false>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else if (acc_mode == 1 && !smooth_v)
    launch_stage3b_v<N, true, 1, false>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else if (acc_mode == 2 && !smooth_v)
    launch_stage3b_v<N, true, 2, false>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else if (acc_mode == 1 && smooth_v)
    launch_stage3b_v<N, true, 1, true>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else if (acc_mode == 2 && smooth_v)
    launch_stage3b_v<N, true, 2, true>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else if (acc_mode == 3 && !smooth_v)
    launch_stage3b_v<N, true, 3, false>(m, nb, uin, kp, acc, kout, bh, ch, nu);
  else
    return gf_set_err(m->ctx, GF_ERR_ARG, "stage3b: unsupported stage form");
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int N>
int launch_rk_out_box(gf_mesh *m, double *const *u, double *const *acc, double *const *kf,
                      double cl) {
  if constexpr (N == 8) {
    if (m->halo8) {
      hipLaunchKernelGGL(k_rk_out_halo8<N>, dim3(xcd_grid(m->nelems)), dim3(64), 0, m->ctx->stream,
                         u[0], u[1], u[2], acc[0], acc[1], acc[2], kf[0], kf[1], kf[2], cl,
                         m->box, m->nelems);
      GF_LAUNCH_CHECK(m->ctx);
      return GF_OK;
    }
  }
  constexpr int NN = N * N * N;
  constexpr int CE = NN >= kThreads ? 1 : kThreads / NN;
  size_t nchunks = (m->nelems + CE - 1) / CE;
  int nb = xcd_grid(std::min<size_t>(nchunks, 8192));
  if (m->ndof < ((size_t)1 << 31))
    hipLaunchKernelGGL((k_rk_out_box<N, int32_t>), dim3(nb), dim3(kThreads), 0, m->ctx->stream,
                       u[0], u[1], u[2], acc[0], acc[1], acc[2], kf[0], kf[1], kf[2], cl, m->box,
                       m->nelems);
  else
    hipLaunchKernelGGL((k_rk_out_box<N, int64_t>), dim3(nb), dim3(kThreads), 0, m->ctx->stream,
                       u[0], u[1], u[2], acc[0], acc[1], acc[2], kf[0], kf[1], kf[2], cl, m->box,
                       m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int N, typename T>
int launch_helm3p(gf_mesh *m, T p, const T *pf, const T *qf, T qs, const T *u, T *uo,
                  const gf::PenD<T> &D, const T *W, const T *edx, const T *ejac) {
  using S = gf::Pen3<N>;
  int nb = xcd_grid((m->nelems + S::EPB - 1) / S::EPB);
  hipLaunchKernelGGL((k_helm3p<N, T>), dim3(nb), dim3(S::THREADS), 0, m->ctx->stream, u, uo, D, W,
                     edx, ejac, p, pf, qf, qs, m->nelems);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

template <int N, typename T>
int launch_cg_apply3p(gf_mesh *m, const CGApply3Args &a) {
  using S = gf::Pen3<N>;
  int nb = xcd_grid((m->nelems + S::EPB - 1) / S::EPB);
  gf::PenD<T> Dp;
  for (int c = 0; c < 3; c++) {
    Dp.R[c] = (const T *)a.DR[c];
    Dp.C[c] = (const T *)a.DC[c];
  }
  hipLaunchKernelGGL((k_cg_apply3p<N, T>), dim3(nb), dim3(S::THREADS), 0, m->ctx->stream,
                     (const T *)a.r, (const T *)a.dinv, (T *)a.w, (T *)a.q, Dp, (const T *)a.W,
                     (const T *)a.edx, (const T *)a.ejac, (const T *)a.qf, (T)a.qs, (const T *)a.mask, a.p,
                     (const double *)a.scal, a.part, a.nelems, a.topo, a.tmask, a.seg_eoff, a.nseg,
                     (const double *)a.sscal, (const T *)a.dtab, a.deq);
  GF_LAUNCH_CHECK(m->ctx);
  return GF_OK;
}

// Number of workgroups (= partial sums) the 3-D column apply launches.
template <int N>
int helm3c_d(gf_mesh *m, double p, const double *pf, const double *qf, double qs, const double *u,
             double *uo) {
  return launch_helm3p<N, double>(m, p, pf, qf, qs, u, uo, pend_d(m), m->W[0], m->edx, m->ejac);
}
template <int N>
int helm3c_f(gf_mesh *m, float p, const float *qf, float qs, const float *u, float *uo) {
  return launch_helm3p<N, float>(m, p, (const float *)nullptr, qf, qs, u, uo, pend_f(m), m->W32,
                                 m->edx32, m->ejac32);
}
template <int N>
int cgapply3c_d(gf_mesh *m, const CGApply3Args &a) { return launch_cg_apply3p<N, double>(m, a); }
template <int N>
int cgapply3c_f(gf_mesh *m, const CGApply3Args &a) { return launch_cg_apply3p<N, float>(m, a); }

// u3 = uin + bh*k ; acc = accsrc + ch*k   (final-stage input of step_b, :315-320)

// k_cg_update on a structured box: q arrives unassembled (straight from the
// apply) and is assembled here by gather (gf_boxdss.hpp), in the GGFX order,
// so no DSS pass runs between the apply and the update.  The multiplicity
// weight is 1/(group size), which equals GGFX::get_imult.  Work is split in
// chunks of whole elements so that an element's box coordinates are formed
// once per chunk (and are wave-uniform when a chunk is one element).
template <typename T, int NORM, int N, typename I>
__global__ __launch_bounds__(kThreads) void k_cg_update_box(
    T *__restrict__ x, const T *__restrict__ w, T *__restrict__ r, const T *__restrict__ qv,
    const T *__restrict__ dinv, const double *__restrict__ mass, gf::BoxTopo topo, double *scal,
    double *__restrict__ part, size_t nelems) {
  constexpr int NN = N * N * N;
  constexpr int CE = NN >= kThreads ? 1 : kThreads / NN;  // elements per chunk
  constexpr int CN = CE * NN;                              // nodes per chunk
  constexpr int QPC = (CN + kThreads - 1) / kThreads;      // nodes per thread per chunk
  // chunks per batch: every load of a batch is issued before its first
  // store, so a batch should keep enough bytes in flight (FP32 takes two)
  constexpr int BC = (sizeof(T) == 4 && QPC <= 2) ? 2 : 1;
  constexpr int NQ = BC * QPC;
  __shared__ double sred[kThreads / 64];
  if (scal[S_DONE] != 0.0) return;
  const T alpha = (T)scal[S_ALPHA];
  double rho = 0.0, nrm = 0.0;
  // each (XCD-renumbered) block walks a contiguous range of chunks
  const size_t nchunks = (nelems + CE - 1) / CE;
  const size_t cpb = (nchunks + gridDim.x - 1) / gridDim.x;
  const size_t c0 = xcd_block() * cpb, c1 = c0 + cpb < nchunks ? c0 + cpb : nchunks;
  // box coordinates of the walk's next element (chunks of one element)
  int ecw[3] = {0, 0, 0};
  if (CE == 1) gf::box_elem_coords(topo, c0, ecw);
  for (size_t c = c0; c < c1; c += BC) {
    I gs[NQ];
    T dv[NQ], xo[NQ], wo[NQ], ro[NQ], qm[NQ][8];
    int cnt[NQ];
    bool ok[NQ];
#pragma unroll
    for (int bc = 0; bc < BC; bc++) {
      const size_t cc = c + bc;
      int ec0[3];
      if (CE == 1) {
        ec0[0] = ecw[0];
        ec0[1] = ecw[1];
        ec0[2] = ecw[2];
        gf::box_elem_next(topo, ecw);
      } else {
        gf::box_elem_coords(topo, cc * CE, ec0);
      }
      const gf::BoxElemDeltas d0 = gf::box_elem_deltas(topo, N, ec0);  // uniform when CE == 1
#pragma unroll
      for (int qq = 0; qq < QPC; qq++) {
        const int s = bc * QPC + qq;
        const int q = qq * kThreads + (int)threadIdx.x;
        const int el = CE == 1 ? 0 : q / NN;
        const int rr = CE == 1 ? q : q - el * NN;
        const size_t e = cc * CE + el;
        ok[s] = cc < c1 && q < CN && e < nelems;
        gf::BoxElemDeltas de = d0;
        if (CE > 1 && el) {
          int ec[3];
          gf::box_elem_coords(topo, e, ec);
          de = gf::box_elem_deltas(topo, N, ec);
        }
        const int i = rr % N, j = (rr / N) % N, k = rr / (N * N);
        // out-of-range slots read node 0 (always valid) and are not stored
        const I g = ok[s] ? (I)(e * NN + rr) : (I)0;
        gs[s] = g;
        I mem[8];
        if (ok[s]) {
          cnt[s] = gf::box_members<I>(de, N, g, i, j, k, mem);
        } else {
          cnt[s] = 1;
#pragma unroll
          for (int m2 = 0; m2 < 8; m2++) mem[m2] = (I)0;
        }
        dv[s] = dinv ? dinv[g] : T(1);
        xo[s] = x[g];
        wo[s] = w[g];
        ro[s] = r[g];
        gf::box_load8m<I, T>(qv, mem, cnt[s], qm[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < NQ; s++) {
      if (!ok[s]) continue;
      const I g = gs[s];
      const T qa = gf::box_sum8z<T>(qm[s], cnt[s]);
      const T xv = (T)1 * xo[s] + alpha * wo[s];
      x[g] = xv;
      const T rv = (T)1 * ro[s] + (-alpha) * qa;
      r[g] = rv;
      const T z = dinv ? dv[s] * rv : rv;
      const double im = gf::box_inv(cnt[s]);
      rho += (double)rv * (double)z * im;
      const double rd = (double)rv;
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

// 16-byte vector of T (4 floats or 2 doubles).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using V = float4;
  static constexpr int R = 4;
  __device__ static float get(const V &v, int n) {
    return n == 0 ? v.x : (n == 1 ? v.y : (n == 2 ? v.z : v.w));
  }
  __device__ static void set(V &v, int n, float a) {
    if (n == 0) v.x = a;
    else if (n == 1) v.y = a;
    else if (n == 2) v.z = a;
    else v.w = a;
  }
};
template <>
struct Vec16<double> {
  using V = double2;
  static constexpr int R = 2;
  __device__ static double get(const V &v, int n) { return n == 0 ? v.x : v.y; }
  __device__ static void set(V &v, int n, double a) {
    if (n == 0) v.x = a;
    else v.y = a;
  }
};

// True when k_cg_update_boxv handles order N in precision T: a node row
// splits into at least two 16-byte runs and a workgroup holds whole elements.
template <typename T, int N>
constexpr bool boxv_ok() {
  return N % Vec16<T>::R == 0 && N > Vec16<T>::R &&
         kThreads % (N * N * N / Vec16<T>::R) == 0;
}

// k_cg_update_box with each lane working a run of R = 16/sizeof(T)
// consecutive nodes along x (one 16-byte load or store per vector), so the
// wave issues a quarter (FP32) or half (FP64) of the memory instructions.
// The group of a node is the product of its partner choices per direction:
// the run's y, z and yz partners are runs of the neighbouring elements at the
// same x offset (16-byte loads), and only the run's first or last node can
// have an x partner (scalar loads).  Members are summed in the order
// (z side, y side, x side), lower side first, where the lower side of a
// direction is the copy with the smaller global index; because the element
// index dominates the global index, that order is the ascending order of the
// group's global indices, GGFX's summation order (ggfx.hpp:346-443), so the
// result equals k_cg_update_box's bit for bit.
template <typename T, int NORM, int N, typename I>
__global__ __launch_bounds__(kThreads) void k_cg_update_boxv(
    T *__restrict__ x, const T *__restrict__ w, T *__restrict__ r, const T *__restrict__ qv,
    const T *__restrict__ dinv, const double *__restrict__ mass, gf::BoxTopo topo, double *scal,
    double *__restrict__ part, size_t nelems, const T *__restrict__ dtab,
    const unsigned char *__restrict__ deq) {
  using VT = Vec16<T>;
  using V = typename VT::V;
  constexpr int R = VT::R;
  constexpr int NN = N * N * N;
  constexpr int RPR = N / R;           // runs per node row
  constexpr int RPE = NN / R;          // runs per element
  constexpr int EPP = kThreads / RPE;  // elements per workgroup pass
  __shared__ double sred[kThreads / 64];
  if (scal[S_DONE] != 0.0) return;
  const T alpha = (T)scal[S_ALPHA];
  double rho = 0.0, nrm = 0.0;
  const int el = (int)threadIdx.x / RPE, rr = (int)threadIdx.x - el * RPE;
  const int ir = rr % RPR, j = (rr / RPR) % N, k = rr / (RPR * N);
  const int i0 = ir * R;
  // the run node that can have an x partner: the row's first or last node
  const int xn = ir == 0 ? 0 : (ir == RPR - 1 ? R - 1 : -1);
  const size_t npass = (nelems + EPP - 1) / EPP;
  const size_t ppb = (npass + gridDim.x - 1) / gridDim.x;
  const size_t c0 = xcd_block() * ppb, c1 = c0 + ppb < npass ? c0 + ppb : npass;
  for (size_t c = c0; c < c1; c++) {
    const size_t e = c * EPP + el;
    if (e >= nelems) continue;
    int ec[3];
    gf::box_elem_coords(topo, e, ec);
    const gf::BoxElemDeltas d = gf::box_elem_deltas(topo, N, ec);
    const I g0 = (I)(e * NN) + (I)(k * N * N + j * N + i0);
    const I dy = j == 0 ? (I)d.lo[1] : (j == N - 1 ? (I)d.hi[1] : (I)0);
    const I dz = k == 0 ? (I)d.lo[2] : (k == N - 1 ? (I)d.hi[2] : (I)0);
    const I dx = xn == 0 ? (I)d.lo[0] : (xn == R - 1 ? (I)d.hi[0] : (I)0);
    const bool py = dy != 0, pz = dz != 0, px = dx != 0;
    // every load of the run before its first store
    const V xo = *(const V *)(x + g0);
    const V wo = *(const V *)(w + g0);
    const V ro = *(const V *)(r + g0);
    V dv;
    if (dinv)  // the Jacobi values from the table where the element's equal it
      dv = (dtab && deq[e]) ? *(const V *)(dtab + (g0 - (I)(e * NN))) : *(const V *)(dinv + g0);
    const V q0 = *(const V *)(qv + g0);
    V qy = {}, qz = {}, qyz = {};
    if (py) qy = *(const V *)(qv + g0 + dy);
    if (pz) qz = *(const V *)(qv + g0 + dz);
    if (py && pz) qyz = *(const V *)(qv + g0 + dy + dz);
    T sx = T(0), sxy = T(0), sxz = T(0), sxyz = T(0);
    if (px) {
      const I gx = g0 + (I)xn + dx;
      sx = qv[gx];
      if (py) sxy = qv[gx + dy];
      if (pz) sxz = qv[gx + dz];
      if (py && pz) sxyz = qv[gx + dy + dz];
    }
    // the partner side is the lower one when its offset is negative
    const bool ysw = dy < 0, zsw = dz < 0, xsw = dx < 0;
    V xv, rv;
#pragma unroll
    for (int n = 0; n < R; n++) {
      // t[side z][side y][side x], side 0 = own copy, 1 = partner
      T t[2][2][2];
      t[0][0][0] = VT::get(q0, n);
      t[0][1][0] = VT::get(qy, n);
      t[1][0][0] = VT::get(qz, n);
      t[1][1][0] = VT::get(qyz, n);
      const bool hx = (n == 0 || n == R - 1) && n == xn && px;
      t[0][0][1] = hx ? sx : T(0);
      t[0][1][1] = hx ? sxy : T(0);
      t[1][0][1] = hx ? sxz : T(0);
      t[1][1][1] = hx ? sxyz : T(0);
      const bool xs = hx && xsw;
      // canonical order: lower side first in z, then y, then x; absent
      // partners contribute +0 (see box_sum8z)
      T sum = T(0);
#pragma unroll
      for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
          for (int cx = 0; cx < 2; cx++) {
            const int sz = cz ^ (int)zsw, sy = cy ^ (int)ysw, sxi = cx ^ (int)xs;
            sum = sum + t[sz][sy][sxi];
          }
      const int cnt = (1 + (int)py) * (1 + (int)pz) * (1 + (int)hx);
      const T qa = cnt == 1 ? t[0][0][0] : sum;
      const T xn1 = (T)1 * VT::get(xo, n) + alpha * VT::get(wo, n);
      VT::set(xv, n, xn1);
      const T rn1 = (T)1 * VT::get(ro, n) + (-alpha) * qa;
      VT::set(rv, n, rn1);
      const T z = dinv ? VT::get(dv, n) * rn1 : rn1;
      const double im = gf::box_inv(cnt);
      rho += (double)rn1 * (double)z * im;
      const double rd = (double)rn1;
      if (NORM == GF_NORM_INF) nrm = fmax(nrm, fabs(rd));
      else if (NORM == GF_NORM_EUC) nrm += rd * rd;
      else if (NORM == GF_NORM_L2) nrm += mass[g0 + n] * rd * rd;
      else if (NORM == GF_NORM_L1) nrm += mass[g0 + n] * fabs(rd);
      else nrm += rd * rd * im;
    }
    *(V *)(x + g0) = xv;
    *(V *)(r + g0) = rv;
  }
  double a = gf::block_reduce<false>(rho, sred);
  double b = NORM == GF_NORM_INF ? gf::block_reduce<true>(nrm, sred)
                                 : gf::block_reduce<false>(nrm, sred);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

template <typename T, int NORM, int N>
int launch_update_box_n(gf_mesh *m, CGVecs<T> &v, double *part) {
  if constexpr (boxv_ok<T, N>()) {
    if (m->ndof < ((size_t)1 << 31))
      hipLaunchKernelGGL((k_cg_update_boxv<T, NORM, N, int32_t>), dim3(kUpdBlocks), dim3(kThreads),
                         0, m->ctx->stream, v.x, v.w, v.r, v.q, v.dinv, m->mass, m->box, m->scal,
                         part, m->nelems, v.dtab, v.deq);
    else
      hipLaunchKernelGGL((k_cg_update_boxv<T, NORM, N, int64_t>), dim3(kUpdBlocks), dim3(kThreads),
                         0, m->ctx->stream, v.x, v.w, v.r, v.q, v.dinv, m->mass, m->box, m->scal,
                         part, m->nelems, v.dtab, v.deq);
    return GF_OK;
  }
  // 32-bit node indices whenever every index of the mesh fits
  if (m->ndof < ((size_t)1 << 31))
    hipLaunchKernelGGL((k_cg_update_box<T, NORM, N, int32_t>), dim3(kUpdBlocks), dim3(kThreads), 0,
                       m->ctx->stream, v.x, v.w, v.r, v.q, v.dinv, m->mass, m->box, m->scal, part,
                       m->nelems);
  else
    hipLaunchKernelGGL((k_cg_update_box<T, NORM, N, int64_t>), dim3(kUpdBlocks), dim3(kThreads), 0,
                       m->ctx->stream, v.x, v.w, v.r, v.q, v.dinv, m->mass, m->box, m->scal, part,
                       m->nelems);
  return GF_OK;
}

}  // namespace gfs

#define GF_K3D_UPD(T, N)                                                              \
  template int launch_update_box_n<T, GF_NORM_INF, N>(gf_mesh *, CGVecs<T> &, double *); \
  template int launch_update_box_n<T, GF_NORM_EUC, N>(gf_mesh *, CGVecs<T> &, double *); \
  template int launch_update_box_n<T, GF_NORM_L2, N>(gf_mesh *, CGVecs<T> &, double *);  \
  template int launch_update_box_n<T, GF_NORM_L1, N>(gf_mesh *, CGVecs<T> &, double *);  \
  template int launch_update_box_n<T, GF_NORM_ASSEMBLED_L2, N>(gf_mesh *, CGVecs<T> &, double *);

#define GF_K3D_INSTANTIATE(N)                                                                \
  template int launch_stage3p<N>(gf_mesh *, const StageArgs &);                              \
  template int launch_stage3b<N>(gf_mesh *, double *const *, double *const *, double *const *, \
                                 double *const *, double, double, double, int, int);         \
  template int launch_rk_out_box<N>(gf_mesh *, double *const *, double *const *,              \
                                    double *const *, double);                                \
  template int helm3c_d<N>(gf_mesh *, double, const double *, const double *, double,         \
                           const double *, double *);                                        \
  template int helm3c_f<N>(gf_mesh *, float, const float *, float, const float *, float *);   \
  template int cgapply3c_d<N>(gf_mesh *, const CGApply3Args &);                              \
  template int cgapply3c_f<N>(gf_mesh *, const CGApply3Args &);                              \
  GF_K3D_UPD(float, N)                                                                       \
  GF_K3D_UPD(double, N)
