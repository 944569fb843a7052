// gf_ggfx_dev.hip -- GGFX coincidence maps of a structured 3-D box built on
// the device.
//
// GGFX::init (src/cdg/comm/ggfx.hpp:167-342) matches nodes by coordinates:
// every node's group is the set of local nodes at the same position.  On a
// box of regular elements in x-fastest linear order that set follows from the
// node's position in the element lattice (gf_boxdss.hpp): in each direction a
// node with index 0 coincides with index N-1 of the previous element and vice
// versa, wrapping where the direction is periodic.  The maps below are formed
// from that rule per node, so no coordinates are compared and the work is one
// pass over the nodes plus an exclusive scan of the group sizes.
//
// The rule reproduces the coordinate matching of the reference only where
// every direction is periodic or none is: on a box with periodic and
// non-periodic directions GGridBox::periodize moves a node on any periodic
// face in every coordinate that lies on a P1 face (ggrid_box.ipp:753-802),
// which joins groups the lattice keeps apart; such boxes are refused
// (GF_ERR_UNSUPPORTED) and keep gf_ggfx_build.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "gf_boxdss.hpp"
#include "gf_internal.hpp"
#include "gf_kernels.hpp"

namespace {

constexpr int kT = 256;

struct NodePos {
  size_t e;
  int i, j, k;
};

__device__ __forceinline__ NodePos node_pos(size_t q, int N) {
  const size_t NN = (size_t)N * N * N;
  NodePos p;
  p.e = q / NN;
  const int r = (int)(q - p.e * NN);
  p.i = r % N;
  p.j = (r / N) % N;
  p.k = r / (N * N);
  return p;
}

// Group size of every node; with lead != NULL also the group size at the
// group's smallest member (the node at which GGFX forms the group) for
// groups of more than one member, else 0, and a 0/1 leader flag.
// The box's nodes are q0 .. q0 + n - 1 of the mesh (q0 = 0 for a single box).
__global__ void k_box_count(gf::BoxTopo t, int N, size_t n, int64_t *__restrict__ cnt,
                            int *__restrict__ lead, int *__restrict__ isl, size_t q0) {
  for (size_t q = blockIdx.x * (size_t)kT + threadIdx.x; q < n; q += (size_t)gridDim.x * kT) {
    const NodePos p = node_pos(q, N);
    int ec[3];
    gf::box_elem_coords(t, p.e, ec);
    const gf::BoxElemDeltas de = gf::box_elem_deltas(t, N, ec);
    int64_t g[8];
    const int c = gf::box_members<int64_t>(de, N, (int64_t)q, p.i, p.j, p.k, g);
    if (cnt) cnt[q0 + q] = c;
    if (lead) {
      const bool l = c > 1 && g[0] == (int64_t)q;
      lead[q0 + q] = l ? c : 0;
      isl[q0 + q] = l ? 1 : 0;
    }
  }
}

// members[offs[q] ..) = the group of node q in ascending order.
__global__ void k_box_fill(gf::BoxTopo t, int N, size_t n, const int64_t *__restrict__ offs,
                           int64_t *__restrict__ members) {
  for (size_t q = blockIdx.x * (size_t)kT + threadIdx.x; q < n; q += (size_t)gridDim.x * kT) {
    const NodePos p = node_pos(q, N);
    int ec[3];
    gf::box_elem_coords(t, p.e, ec);
    const gf::BoxElemDeltas de = gf::box_elem_deltas(t, N, ec);
    int64_t g[8];
    const int c = gf::box_members<int64_t>(de, N, (int64_t)q, p.i, p.j, p.k, g);
    const int64_t o = offs[q];
    for (int s = 0; s < c; s++) members[o + s] = g[s];
  }
}

// The mesh's compact form: groups of more than one member, each once, in
// the order of their smallest member; goff[gi] = start of group gi in gmem.
__global__ void k_box_groups(gf::BoxTopo t, int N, size_t n, const int *__restrict__ loff,
                             const int *__restrict__ gidx, const int *__restrict__ isl,
                             int *__restrict__ goff, int *__restrict__ gmem, size_t q0) {
  for (size_t q = blockIdx.x * (size_t)kT + threadIdx.x; q < n; q += (size_t)gridDim.x * kT) {
    if (!isl[q0 + q]) continue;
    const NodePos p = node_pos(q, N);
    int ec[3];
    gf::box_elem_coords(t, p.e, ec);
    const gf::BoxElemDeltas de = gf::box_elem_deltas(t, N, ec);
    int64_t g[8];
    const int c = gf::box_members<int64_t>(de, N, (int64_t)q, p.i, p.j, p.k, g);
    const int o = loff[q0 + q];
    goff[gidx[q0 + q]] = o;
    for (int s = 0; s < c; s++) gmem[o + s] = (int)(q0 + g[s]);
  }
}

int grid_for(size_t n) {
  size_t b = (n + kT - 1) / kT;
  return (int)std::min<size_t>(std::max<size_t>(b, 1), 65536);
}

template <typename T>
int exclusive_scan(gf_ctx *ctx, const T *in, T *out, size_t n) {
  size_t tb = 0;
  GF_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, in, out, T(0), n, rocprim::plus<T>(),
                                      ctx->stream));
  void *tmp = nullptr;
  GF_HIP(ctx, hipMalloc(&tmp, tb > 0 ? tb : 1));
  hipError_t e = rocprim::exclusive_scan(tmp, tb, in, out, T(0), n, rocprim::plus<T>(),
                                         ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(tmp);
  return e == hipSuccess ? GF_OK : gf_check_hip(ctx, e, "exclusive scan");
}

int box_topo(gf_ctx *ctx, int N, const int *ne, const int *periodic, gf::BoxTopo *t,
             size_t *ndof) {
  if (!ctx || !ne || !periodic || N < 2 || N > GF_MAX_N) return GF_ERR_ARG;
  size_t nel = 1;
  int np = 0;
  for (int d = 0; d < 3; d++) {
    if (ne[d] < 1) return gf_set_err(ctx, GF_ERR_ARG, "box: element counts must be positive");
    t->ne[d] = ne[d];
    t->per[d] = periodic[d] ? 1 : 0;
    np += t->per[d];
    nel *= (size_t)ne[d];
  }
  for (int f = 0; f < 6; f++) t->fmask[f] = 0;
  if (np != 0 && np != 3)
    return gf_set_err(ctx, GF_ERR_UNSUPPORTED,
                      "box GGFX on the device: every direction periodic or none");
  *ndof = nel * (size_t)N * N * N;
  return GF_OK;
}

}  // namespace

// Compact device groups of a mesh made of nbox boxes in sequence (box b:
// elements ne[3b..3b+2], periodic flags per[3b..3b+2]), as gf_mesh_create
// builds them when the descriptor asks for device-built maps: the groups of
// each box, in the order of their smallest member; m->goff / m->gmem are
// allocated here.  A single box also sets m->box.
int gf_box_groups_device(gf_mesh *m, int nbox, const int *ne, const int *periodic) {
  gf_ctx *ctx = m->ctx;
  if (nbox < 1) return GF_ERR_ARG;
  std::vector<gf::BoxTopo> topo(nbox);
  std::vector<size_t> q0(nbox + 1, 0);
  for (int b = 0; b < nbox; b++) {
    size_t nb = 0;
    int rc = box_topo(ctx, m->Niso, ne + 3 * b, periodic + 3 * b, &topo[b], &nb);
    if (rc) return rc;
    q0[b + 1] = q0[b] + nb;
  }
  const size_t n = q0[nbox];
  if (n != m->ndof) return gf_set_err(ctx, GF_ERR_ARG, "box GGFX: element counts do not match");
  if (n >= ((size_t)1 << 31)) return gf_set_err(ctx, GF_ERR_UNSUPPORTED, "box GGFX: > 2^31 nodes");
  int rc = GF_OK;
  int *lead = nullptr, *isl = nullptr, *loff = nullptr, *gidx = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(lead);
    (void)hipFree(isl);
    (void)hipFree(loff);
    (void)hipFree(gidx);
  };
  if (hipMalloc((void **)&lead, (n + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&isl, (n + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&loff, (n + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&gidx, (n + 1) * sizeof(int)) != hipSuccess) {
    cleanup();
    return gf_set_err(ctx, GF_ERR_ALLOC, "box GGFX workspace");
  }
  for (int b = 0; b < nbox; b++) {
    const size_t nb = q0[b + 1] - q0[b];
    hipLaunchKernelGGL(k_box_count, dim3(grid_for(nb)), dim3(kT), 0, ctx->stream, topo[b], m->Niso,
                       nb, (int64_t *)nullptr, lead, isl, q0[b]);
  }
  // one more (zero) entry so the scans also give the totals
  hipError_t e = hipMemsetAsync(lead + n, 0, sizeof(int), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(isl + n, 0, sizeof(int), ctx->stream);
  if (e != hipSuccess) {
    cleanup();
    return gf_check_hip(ctx, e, "box GGFX");
  }
  rc = exclusive_scan<int>(ctx, lead, loff, n + 1);
  if (!rc) rc = exclusive_scan<int>(ctx, isl, gidx, n + 1);
  int tot[2] = {0, 0};
  if (!rc) {
    e = hipMemcpy(&tot[0], loff + n, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&tot[1], gidx + n, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = gf_check_hip(ctx, e, "box GGFX");
  }
  if (!rc) {
    const int nmem = tot[0], ng = tot[1];
    if (hipMalloc((void **)&m->goff, (ng + 1) * sizeof(int) + 16) != hipSuccess ||
        hipMalloc((void **)&m->gmem, (std::max(nmem, 1)) * sizeof(int) + 16) != hipSuccess) {
      rc = gf_set_err(ctx, GF_ERR_ALLOC, "box GGFX groups");
    } else {
      for (int b = 0; b < nbox; b++) {
        const size_t nb = q0[b + 1] - q0[b];
        hipLaunchKernelGGL(k_box_groups, dim3(grid_for(nb)), dim3(kT), 0, ctx->stream, topo[b],
                           m->Niso, nb, loff, gidx, isl, m->goff, m->gmem, q0[b]);
      }
      e = hipMemcpyAsync(m->goff + ng, &tot[0], sizeof(int), hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = gf_check_hip(ctx, e, "box GGFX groups");
      m->ngroups = ng;
    }
  }
  cleanup();
  if (rc) return rc;
  // largest group: two members per direction in which nodes have partners
  int mg = 1;
  for (int b = 0; b < nbox; b++) {
    int g = 1;
    for (int d = 0; d < 3; d++) g *= (topo[b].per[d] || topo[b].ne[d] > 1) ? 2 : 1;
    mg = std::max(mg, g);
  }
  m->maxgrp = m->ngroups > 0 ? mg : 1;
  if (nbox == 1) m->box = topo[0];
  return GF_OK;
}

extern "C" int gf_box_ggfx_device(gf_ctx *ctx, int N, const int *ne, const int *periodic,
                                  int64_t *offs, int64_t *members, size_t *nmembers) {
  gf::BoxTopo t;
  size_t n = 0;
  int rc = box_topo(ctx, N, ne, periodic, &t, &n);
  if (rc) return rc;
  if (!offs || !nmembers) return GF_ERR_ARG;
  GF_HIP(ctx, hipSetDevice(ctx->device));
  int64_t *cnt = nullptr;
  GF_HIP(ctx, hipMalloc((void **)&cnt, (n + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(k_box_count, dim3(grid_for(n)), dim3(kT), 0, ctx->stream, t, N, n, cnt,
                     (int *)nullptr, (int *)nullptr, (size_t)0);
  hipError_t e = hipMemsetAsync(cnt + n, 0, sizeof(int64_t), ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(cnt);
    return gf_check_hip(ctx, e, "box GGFX");
  }
  rc = exclusive_scan<int64_t>(ctx, cnt, offs, n + 1);
  (void)hipFree(cnt);
  if (rc) return rc;
  int64_t tot = 0;
  GF_HIP(ctx, hipMemcpy(&tot, offs + n, sizeof(int64_t), hipMemcpyDeviceToHost));
  *nmembers = (size_t)tot;
  if (members) {
    hipLaunchKernelGGL(k_box_fill, dim3(grid_for(n)), dim3(kT), 0, ctx->stream, t, N, n, offs,
                       members);
    GF_LAUNCH_CHECK(ctx);
    GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GF_OK;
}
