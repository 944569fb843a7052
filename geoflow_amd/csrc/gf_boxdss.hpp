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

// ---------------------------------------------------------------------------
// Element halo in LDS for a gather Smooth at N = 8 (one wave of 64 lanes per
// element, lane = a + 8 b owning the column (a, b, .)).  Instead of each node
// loading its group members from global memory -- loads under per-lane
// conditions, each waited for in turn -- the wave loads the element's own
// values and every partner value it can need with 17 unconditional loads
// per lane, all in flight together, and the group sums read LDS.
// Layout (doubles):
//   own values       [0, 512)     i + 8 j + 64 k
//   faces            [512, 896)   512 + 64 f + p + 8 q, f = 2 d + s (d the
//                                 face direction, s = 1 for the upper face),
//                                 (p, q) the node's other two indices in
//                                 increasing direction order
//   edges            [896, 992)   896 + 8 e + r, e = 4 pd + 2 s1 + s2 for the
//                                 direction pair pd (0: xy, 1: xz, 2: yz) with
//                                 sides s1, s2, r the node's third index
//   corners          [992, 1000)  992 + sx + 2 sy + 4 sz
//   zero             1000         +0, read for absent partners
// A halo slot holds the value at the partner node across that face, edge
// or corner (the offsets of box_elem_deltas added up); where the element has
// no partner the offset is 0, the slot holds an own value and is never read.
// ---------------------------------------------------------------------------
constexpr int kHalo8 = 1001;

// f(std::integral_constant<int, B>), ..., f(std::integral_constant<int, E-1>):
// a loop whose index is a compile-time constant in the body.
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}
constexpr int kHalo8Zero = 1000;

__device__ __forceinline__ int64_t halo8_delta(const BoxElemDeltas &de, int d, int s) {
  return s ? de.hi[d] : de.lo[d];
}

// Offset from the element's first node read by load q (0..8) of lane `lane`
// and the LDS slot it fills (-1: the lane has nothing to load in this round;
// it reads the element's first node instead, so every load is
// unconditional and in bounds).
__device__ __forceinline__ int64_t halo8_src(const BoxElemDeltas &de, int lane, int q, int *slot) {
  if (q < 6) {  // faces: the node's own offset is a lane function per face
    const int d = q >> 1, s = q & 1, p = lane & 7, r = lane >> 3;
    const int own = d == 0 ? 8 * lane + 7 * s : (d == 1 ? p + 64 * r + 56 * s : lane + 448 * s);
    *slot = 512 + 64 * q + lane;
    return own + halo8_delta(de, d, s);
  }
  if (q < 8) {  // edges: 12 x 8 values over two rounds
    const int e = (q - 6) * 8 + (lane >> 3), r = lane & 7;
    if (e >= 12) {
      *slot = -1;
      return 0;
    }
    // pair pd: 0 (x, y), 1 (x, z), 2 (y, z), sides s1, s2; selects only,
    // so no array is indexed by a lane-dependent value
    const int pd = e >> 2, s1 = (e >> 1) & 1, s2 = e & 1;
    const int e1 = s1 ? 7 : 0, e2 = s2 ? 7 : 0;
    const int c0 = pd == 2 ? r : e1;
    const int c1 = pd == 0 ? e2 : (pd == 1 ? r : e1);
    const int c2 = pd == 0 ? r : e2;
    const int64_t dx = halo8_delta(de, 0, s1), dy1 = halo8_delta(de, 1, s1),
                  dy2 = halo8_delta(de, 1, s2), dz = halo8_delta(de, 2, s2);
    *slot = 896 + 8 * e + r;
    return c0 + 8 * c1 + 64 * c2 + (pd == 0 ? dx + dy2 : (pd == 1 ? dx + dz : dy1 + dz));
  }
  if (lane >= 8) {  // corners
    *slot = -1;
    return 0;
  }
  const int sx = lane & 1, sy = (lane >> 1) & 1, sz = lane >> 2;
  *slot = 992 + lane;
  return 7 * sx + 56 * sy + 448 * sz + halo8_delta(de, 0, sx) + halo8_delta(de, 1, sy) +
         halo8_delta(de, 2, sz);
}

// The partner values of field u for the element at ebase: 9 loads per lane
// into registers (halo8_src), then their LDS slots in the halo hl (kHalo8
// values) and the zero slot; the element's own values, hl[0, 512), are
// copied separately (one 4 KB block in the element's layout).
template <typename T>
struct Halo8Rim {
  T hv[9];
  int slot[9];
  __device__ __forceinline__ void issue(const T *__restrict__ u, const BoxElemDeltas &de,
                                        int64_t ebase, int lane) {
#pragma unroll
    for (int q = 0; q < 9; q++) hv[q] = u[ebase + halo8_src(de, lane, q, &slot[q])];
  }
  __device__ __forceinline__ void store(T *hl, int lane) const {
#pragma unroll
    for (int q = 0; q < 9; q++)
      if (slot[q] >= 0) hl[slot[q]] = hv[q];
    if (lane == 0) hl[kHalo8Zero] = T(0);
  }
};

// True when halo8_order's canonical order is GGFX's on this box at N = 8:
// every partner offset in x smaller than every one in y, and the x and y ones
// together smaller than every z one (superincreasing in x, y, z).  An offset
// is the element stride less 7 node strides between neighbours, (ne - 1)
// element strides plus 7 node strides across a periodic wrap, and 7 node
// strides within an element that is its own periodic neighbour (ne = 1).
__host__ __device__ __forceinline__ bool box_halo8_ok(const BoxTopo &t) {
  const int64_t NN = 512, ns[3] = {1, 8, 64};
  const int64_t es[3] = {NN, (int64_t)t.ne[0] * NN, (int64_t)t.ne[0] * t.ne[1] * NN};
  int64_t below = 0;  // sum of the largest offsets of the lower directions
  bool ok = true;
  for (int d = 0; d < 3; d++) {
    int64_t mn, mx;
    if (t.ne[d] == 1) {
      if (!t.per[d]) continue;  // no partners in this direction
      mn = mx = 7 * ns[d];
    } else {
      mn = es[d] - 7 * ns[d];
      mx = t.per[d] ? (int64_t)(t.ne[d] - 1) * es[d] + 7 * ns[d] : mn;
    }
    ok = ok && below < mn;
    below += mx;
  }
  return ok;
}

// Per-lane constants of the group sums of the column (i, j, .): which x and
// y partners exist, whether they lie below, and the k-independent parts of
// their halo addresses (the zero slot where a partner is absent).
struct Halo8Col {
  int i, j, xon, yon, nx, ny;
  int ax0, ay0, axy0;  // + k * (8 xon), + k * (8 yon), + k * (xon & yon)
  __device__ __forceinline__ Halo8Col(const BoxElemDeltas &de, int i_, int j_) : i(i_), j(j_) {
    const int sx = i == 7, sy = j == 7;
    const int64_t dx = i == 0 ? de.lo[0] : (i == 7 ? de.hi[0] : 0);
    const int64_t dy = j == 0 ? de.lo[1] : (j == 7 ? de.hi[1] : 0);
    xon = dx != 0;
    yon = dy != 0;
    nx = dx < 0;
    ny = dy < 0;
    ax0 = xon ? 512 + 64 * sx + j : kHalo8Zero;
    ay0 = yon ? 640 + 64 * sy + i : kHalo8Zero;
    axy0 = (xon & yon) ? 896 + 16 * sx + 8 * sy : kHalo8Zero;
  }
};

// The group of node (i, j, K) as halo addresses in GGFX order: the partner
// offsets' magnitudes are superincreasing x < y < z for N = 8 (an element
// has 512 nodes, more than any in-element offset), so ascending global
// index is the canonical order -- lower side first in z, then in y, then in
// x -- over the 2 x 2 (x 2 on the z faces) crossings, an absent partner's
// terms reading the zero slot, which adds +0 and leaves every sum as it is
// (box_sum8z).  Returns the count of the group.
template <int K>
__device__ __forceinline__ int halo8_order(const Halo8Col &c, const BoxElemDeltas &de,
                                           int (&a)[8]) {
  const int own = c.i + 8 * c.j + 64 * K;
  const int ax = c.ax0 + K * 8 * c.xon, ay = c.ay0 + K * 8 * c.yon;
  const int axy = c.axy0 + K * (c.xon & c.yon);
  // within a y row, lower x first; then lower y row first
  const int l0 = c.nx ? ax : own, u0 = c.nx ? own : ax;
  const int l1 = c.nx ? axy : ay, u1 = c.nx ? ay : axy;
  int p[4] = {c.ny ? l1 : l0, c.ny ? u1 : u0, c.ny ? l0 : l1, c.ny ? u0 : u1};
  int zon = 0, nz = 0;
  int q[4] = {kHalo8Zero, kHalo8Zero, kHalo8Zero, kHalo8Zero};
  if (K == 0 || K == 7) {
    const int64_t dz = K == 0 ? de.lo[2] : de.hi[2];
    zon = dz != 0;
    nz = dz < 0;
    const int sz = K == 7, sx = c.i == 7, sy = c.j == 7;
    const int bz = zon ? 768 + 64 * sz + c.i + 8 * c.j : kHalo8Zero;
    const int bxz = (zon & c.xon) ? 928 + 16 * sx + 8 * sz + c.j : kHalo8Zero;
    const int byz = (zon & c.yon) ? 960 + 16 * sy + 8 * sz + c.i : kHalo8Zero;
    const int bxyz = (zon & c.xon & c.yon) ? 992 + sx + 2 * sy + 4 * sz : kHalo8Zero;
    const int m0 = c.nx ? bxz : bz, n0 = c.nx ? bz : bxz;
    const int m1 = c.nx ? bxyz : byz, n1 = c.nx ? byz : bxyz;
    q[0] = c.ny ? m1 : m0;
    q[1] = c.ny ? n1 : n0;
    q[2] = c.ny ? m0 : m1;
    q[3] = c.ny ? n0 : n1;
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    a[m] = nz ? q[m] : p[m];
    a[4 + m] = nz ? p[m] : q[m];
  }
  return (1 + c.xon) * (1 + c.yon) * (1 + zon);
}

}  // namespace gf
