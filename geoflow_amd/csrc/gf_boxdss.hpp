// gf_boxdss.hpp -- direct stiffness summation by gather on a structured box.
//
// On a box grid (elements in x-fastest linear order, x/y/z element counts
// ne[], optional periodicity per direction) the GGFX coincidence group of a
// node follows from its position alone: in each direction a node with index
// 0 also lives at index N-1 of the previous element, and one with index N-1
// at index 0 of the next element (wrapping around when that direction is
// periodic).  The group is the product of these per-direction choices, at
// most 8 nodes.  A consumer kernel can therefore assemble a local vector on
// the fly ("gather"): every copy of a shared node reads all members of its
// group and sums them in ascending global index, which is the order in which
// GGFX::doOp accumulates a group (std::accumulate from 0 over the sorted
// members, src/cdg/comm/ggfx.hpp:346-443).  All copies compute the same sum,
// so the result stays continuous, and no separate DSS pass or index array is
// needed.  gf_mesh_create enables this only after checking on the host that
// the groups it produces equal the GGFX groups of the mesh.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gf {

struct BoxTopo {
  int ne[3];     // elements per direction
  int per[3];    // 1 if the direction is periodic
  int fmask[6];  // 1 if the nodes of face x0,x1,y0,y1,z0,z1 are masked (Dirichlet-type)
};

// Boundary mask of node (i,j,k) (N per direction) of the element at box
// coordinates ec[]: 0 on a masked domain face, else 1 -- the mask of
// GGrid::init_bc_info for a box whose face types are fmask[].
__host__ __device__ __forceinline__ bool box_masked(const BoxTopo &t, int N, const int (&ec)[3],
                                                    int i, int j, int k) {
  const int nc[3] = {i, j, k};
  bool mk = false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    mk = mk || (nc[d] == 0 && ec[d] == 0 && t.fmask[2 * d]);
    mk = mk || (nc[d] == N - 1 && ec[d] == t.ne[d] - 1 && t.fmask[2 * d + 1]);
  }
  return mk;
}

// Box coordinates (ex, ey, ez) of element e (32-bit division whenever the
// operands fit, which avoids the long 64-bit division sequence).
__host__ __device__ __forceinline__ void box_elem_coords(const BoxTopo &t, size_t e, int (&ec)[3]) {
  const size_t nxy = (size_t)t.ne[0] * t.ne[1];
  if (e < ((size_t)1 << 32) && nxy < ((size_t)1 << 32)) {
    const uint32_t e32 = (uint32_t)e, nx = (uint32_t)t.ne[0], n2 = (uint32_t)nxy;
    const uint32_t z = e32 / n2, rem = e32 - z * n2, y = rem / nx;
    ec[2] = (int)z;
    ec[1] = (int)y;
    ec[0] = (int)(rem - y * nx);
    return;
  }
  ec[2] = (int)(e / nxy);
  const size_t rem = e - (size_t)ec[2] * nxy;
  ec[1] = (int)(rem / (size_t)t.ne[0]);
  ec[0] = (int)(rem - (size_t)ec[1] * t.ne[0]);
}

// Advances box coordinates to those of the next element in linear order.
__host__ __device__ __forceinline__ void box_elem_next(const BoxTopo &t, int (&ec)[3]) {
  if (++ec[0] == t.ne[0]) {
    ec[0] = 0;
    if (++ec[1] == t.ne[1]) {
      ec[1] = 0;
      ++ec[2];
    }
  }
}

// Members of the group of node (i,j,k) (N nodes per direction) of the element
// at box coordinates ec[], whose own global index is gself; written to g[] in
// ascending order, g[m] = SIZE_MAX for m >= count; returns the count.
//
// In direction d the partner differs from the node by the index offset
// delta_d (0 if there is none).  Global indices are e*N^3 + k*N^2 + j*N + i
// with e = ex + nx*(ey + ny*ez), i.e. lexicographic in (ez,ey,ex,k,j,i), so
// the non-zero |delta_d| sorted in decreasing order are superincreasing (each
// exceeds the sum of the smaller ones).  With lo the smallest member, member
// m = lo + sum_s bit_s(m) |delta|_(s) therefore increases with m.
__host__ __device__ __forceinline__ int box_group(const BoxTopo &t, int N, const int (&ec)[3],
                                                  size_t gself, int i, int j, int k,
                                                  size_t (&g)[8]) {
  const int nc[3] = {i, j, k};
  const int64_t NN = (int64_t)N * N * N;
  const int64_t es[3] = {NN, (int64_t)t.ne[0] * NN, (int64_t)t.ne[0] * t.ne[1] * NN};
  const int64_t ns[3] = {1, N, (int64_t)N * N};
  int64_t lo = (int64_t)gself;
  int64_t a[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int64_t dl = 0;
    if (nc[d] == 0 && (ec[d] > 0 || t.per[d])) {
      const int oe = ec[d] > 0 ? ec[d] - 1 : t.ne[d] - 1;
      dl = (int64_t)(oe - ec[d]) * es[d] + (int64_t)(N - 1) * ns[d];
    } else if (nc[d] == N - 1 && (ec[d] < t.ne[d] - 1 || t.per[d])) {
      const int oe = ec[d] < t.ne[d] - 1 ? ec[d] + 1 : 0;
      dl = (int64_t)(oe - ec[d]) * es[d] - (int64_t)(N - 1) * ns[d];
    }
    lo += dl < 0 ? dl : 0;
    a[d] = dl < 0 ? -dl : dl;
  }
  // decreasing magnitudes (absent directions, 0, last)
  auto cs = [](int64_t &x, int64_t &y) {
    const int64_t h = x > y ? x : y, l = x > y ? y : x;
    x = h;
    y = l;
  };
  cs(a[0], a[1]);
  cs(a[1], a[2]);
  cs(a[0], a[1]);
  const int na = (a[0] != 0) + (a[1] != 0) + (a[2] != 0);
  const int cnt = 1 << na;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    int64_t v = lo;
    if (na >= 1) v += ((m >> (na - 1)) & 1) ? a[0] : 0;
    if (na >= 2) v += ((m >> (na - 2)) & 1) ? a[1] : 0;
    if (na >= 3) v += (m & 1) ? a[2] : 0;
    g[m] = m < cnt ? (size_t)v : SIZE_MAX;
  }
  return cnt;
}

// Partner offsets of one element: in direction d a node with index 0 has its
// partner at offset lo[d] (index N-1 of the previous element), a node with
// index N-1 at hi[d] (index 0 of the next element); 0 where there is none.
// They depend only on the element, so kernels form them once per element.
struct BoxElemDeltas {
  int64_t lo[3], hi[3];
};

__host__ __device__ __forceinline__ BoxElemDeltas box_elem_deltas(const BoxTopo &t, int N,
                                                                  const int (&ec)[3]) {
  const int64_t NN = (int64_t)N * N * N;
  const int64_t es[3] = {NN, (int64_t)t.ne[0] * NN, (int64_t)t.ne[0] * t.ne[1] * NN};
  const int64_t ns[3] = {1, N, (int64_t)N * N};
  BoxElemDeltas d;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    d.lo[c] = 0;
    d.hi[c] = 0;
    if (ec[c] > 0 || t.per[c]) {
      const int oe = ec[c] > 0 ? ec[c] - 1 : t.ne[c] - 1;
      d.lo[c] = (int64_t)(oe - ec[c]) * es[c] + (int64_t)(N - 1) * ns[c];
    }
    if (ec[c] < t.ne[c] - 1 || t.per[c]) {
      const int oe = ec[c] < t.ne[c] - 1 ? ec[c] + 1 : 0;
      d.hi[c] = (int64_t)(oe - ec[c]) * es[c] - (int64_t)(N - 1) * ns[c];
    }
  }
  return d;
}

// Members of the group of node (i,j,k) with global index gself, from the
// element's partner offsets; ascending in g[0..count), count returned, and
// *own the slot of gself itself.  Same construction as box_group
// (superincreasing magnitudes), in index type I (int32_t when every index of
// the mesh fits, else int64_t), written without branches or per-slot
// selects: the non-zero magnitudes in ascending order b0 <= b1 <= b2 (0 where
// absent) weight bits 0, 1, 2 of the slot number, so slot m is
// lo + (m&1) b0 + (m&2) b1 + (m&4) b2, and the node's own slot has the bits of
// the directions whose partner lies below it.  Slots past the count repeat
// members (their weights are 0) and are never read.
template <typename I>
__host__ __device__ __forceinline__ int box_members_own(const BoxElemDeltas &d, int N, I gself,
                                                        int i, int j, int k, I (&g)[8], int *own) {
  static_assert(std::is_signed<I>::value, "partner offsets are signed");
  using U = typename std::make_unsigned<I>::type;
  const int nc[3] = {i, j, k};
  I lo = gself;
  U key[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const I dl = nc[c] == 0 ? (I)d.lo[c] : (nc[c] == N - 1 ? (I)d.hi[c] : (I)0);
    lo += dl < 0 ? dl : (I)0;
    // 2 magnitude + (partner below) - 2 as unsigned: an absent direction
    // (magnitude 0) wraps to the largest key and sorts last; magnitudes are
    // distinct, so the sign bit never decides the order
    key[c] = ((((U)(dl < 0 ? -dl : dl)) << 1) | (U)(dl < 0)) - (U)2;
  }
  auto cs = [](U &x, U &y) {
    const U l = x < y ? x : y, h = x < y ? y : x;
    x = l;
    y = h;
  };
  cs(key[0], key[1]);
  cs(key[1], key[2]);
  cs(key[0], key[1]);
  const I b0 = (I)((key[0] + (U)2) >> 1), b1 = (I)((key[1] + (U)2) >> 1),
          b2 = (I)((key[2] + (U)2) >> 1);
  *own = (int)(key[0] & 1) | ((int)(key[1] & 1) << 1) | ((int)(key[2] & 1) << 2);
  const int na = (b0 != 0) + (b1 != 0) + (b2 != 0);
  const I l1 = lo + b1;
  g[0] = lo;
  g[1] = lo + b0;
  g[2] = l1;
  g[3] = l1 + b0;
  g[4] = lo + b2;
  g[5] = g[4] + b0;
  g[6] = l1 + b2;
  g[7] = g[6] + b0;
  return 1 << na;
}

template <typename I>
__host__ __device__ __forceinline__ int box_members(const BoxElemDeltas &d, int N, I gself, int i,
                                                    int j, int k, I (&g)[8]) {
  int own;
  return box_members_own<I>(d, N, gself, i, j, k, g, &own);
}

// 1/count for a box group (count 1, 2, 4 or 8): exact, so x * box_inv(n)
// equals x / n bit for bit and no division is needed.
__host__ __device__ __forceinline__ double box_inv(int n) {
  return n == 1 ? 1.0 : (n == 2 ? 0.5 : (n == 4 ? 0.25 : 0.125));
}

// True when node (i,j,k) can have partners (lies on its element's boundary).
template <int N>
__device__ __forceinline__ bool on_elem_boundary(int i, int j, int k) {
  return i == 0 || i == N - 1 || j == 0 || j == N - 1 || k == 0 || k == N - 1;
}

// Loads u at the 8 member slots of a group (slots past the count re-read the
// node itself), unconditionally: the loads are independent and issue back to
// back, instead of one conditional, waited-for load per member.
template <typename I, typename T>
__device__ __forceinline__ void box_load8(const T *__restrict__ u, const I (&g)[8], int n,
                                          I gself, T (&v)[8]) {
#pragma unroll
  for (int m = 0; m < 8; m++) v[m] = u[m < n ? g[m] : gself];
}

// As box_load8, but absent slots are not read (their lanes are masked off
// and the slot is set to 0), which saves address and cache work when most
// nodes of a wave have small groups.  Slot 0 (the smallest member) is always
// read.
template <typename I, typename T>
__device__ __forceinline__ void box_load8m(const T *__restrict__ u, const I (&g)[8], int n,
                                           T (&v)[8]) {
  v[0] = u[g[0]];
#pragma unroll
  for (int m = 1; m < 8; m++) {
    T t = T(0);
    if (m < n) t = u[g[m]];
    v[m] = t;
  }
}

// 0 + v[0] + v[1] + ... + v[n-1] in that order; n == 1 returns v[0]
// unchanged (singleton groups are not reduced).
template <typename T>
__device__ __forceinline__ T box_sum8(const T (&v)[8], int n) {
  T s = T(0);
#pragma unroll
  for (int m = 0; m < 8; m++) s = m < n ? s + v[m] : s;
  return n == 1 ? v[0] : s;
}

// box_sum8 for slots filled by box_load8m (absent slots hold +0): adding
// them changes nothing, because a sum that starts from +0 is never -0 (in
// round-to-nearest x + y is -0 only when both are -0), and s + (+0) == s for
// every other s.  Same result bit for bit, without a select per slot.
template <typename T>
__device__ __forceinline__ T box_sum8z(const T (&v)[8], int n) {
  T s = T(0);
#pragma unroll
  for (int m = 0; m < 8; m++) s = s + v[m];
  return n == 1 ? v[0] : s;
}

// GGFX Sum of u at node (i,j,k) (global index gself) from the element's
// partner offsets, over the ascending group.  *cnt_out receives the group
// size.
template <int N, typename I, typename T>
__device__ __forceinline__ T box_gather_sum(const BoxElemDeltas &d, const T *__restrict__ u,
                                            I gself, int i, int j, int k, int *cnt_out) {
  I g[8];
  const int n = box_members<I>(d, N, gself, i, j, k, g);
  *cnt_out = n;
  T v[8];
  box_load8<I, T>(u, g, n, gself, v);
  return box_sum8<T>(v, n);
}

}  // namespace gf
