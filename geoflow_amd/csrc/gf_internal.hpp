// gf_internal.hpp -- internal structures shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "geoflow_hip.h"
#include "gf_boxdss.hpp"

#define GF_MAX_N 12  // largest isotropic node count per direction with fused kernels

struct gf_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // device address of the file-scope scratch array of gf_ops.hip on this
  // context's device (hipGetSymbolAddress at gf_ctx_create)
  double *scratch = nullptr;
};

// Reduction slot layout on the device (see gf_reduce.hpp).
#define GF_RED_MAX_BLOCKS 4096

struct gf_mesh_s {
  gf_ctx *ctx = nullptr;
  int dim = 3, gtype = 0;
  int N[3] = {1, 1, 1};
  int Niso = 0;  // isotropic N, 0 if anisotropic
  // dconst: 3-D isotropic mesh whose D is bit-identical to gf::GllD<Niso>
  // (gf_dtab.hpp); the pencil kernels, which take D as compile-time
  // constants, run only then.
  bool dconst = false;
  size_t nelems = 0, nn = 0, ndof = 0;
  // basis (device)
  double *D[3] = {nullptr, nullptr, nullptr};
  double *W[3] = {nullptr, nullptr, nullptr};
  float *D32 = nullptr, *W32 = nullptr;  // isotropic FP32 copies
  // Per-direction row-major / column-major copies of D for the pencil kernels
  double *pDR[3] = {nullptr, nullptr, nullptr}, *pDC[3] = {nullptr, nullptr, nullptr};
  float *pDR32[3] = {nullptr, nullptr, nullptr}, *pDC32[3] = {nullptr, nullptr, nullptr};
  // geometry (device)
  double *edx = nullptr;   // nelems*dim
  double *ejac = nullptr;  // nelems
  float *edx32 = nullptr, *ejac32 = nullptr;
  double *ndx[9] = {nullptr};  // embedded: dXi_i/dX_j at [i*3+j]
  double *njac = nullptr;
  // node metric G: 2-D G00, G01, G11; 3-D deformed G00, G01, G11, G02, G12, G22
  double *G[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double *xyz[3] = {nullptr, nullptr, nullptr};
  double *mask = nullptr, *mass = nullptr, *imass = nullptr, *imult = nullptr;
  // The inverse mass of one element when it is the same, bit for bit, at
  // every element (uniform periodic boxes): nn values the RK stage kernels
  // read in place of imass (null otherwise).
  double *imass_el = nullptr;
  // Structured-box CG with the Jacobi preconditioner: the Jacobi diagonal
  // inverse of element dref (the box's centre) in FP64 and FP32, and per
  // element whether its values equal those bit for bit (set per solve); the
  // CG kernels then read the table for such elements.
  double *dtab = nullptr;
  float *dtab32 = nullptr;
  unsigned char *deq = nullptr;
  size_t dref = 0;
  float *mask32 = nullptr, *imult32 = nullptr;
  bool has_mask = false;
  // DSS: coincidence groups with more than one member (CSR, ascending members)
  int ngroups = 0;
  int *goff = nullptr;
  int *gmem = nullptr;
  int maxgrp = 1;
  // Structured box: when has_box, the groups above are exactly those of
  // gf::box_group for `box`, and fused kernels may assemble by gather.
  gf::BoxTopo box = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0, 0, 0, 0}};
  bool has_box = false;
  // has_box_mask: the boundary mask equals gf::box_masked for `box`, so
  // kernels may form it from the node position instead of reading it.
  bool has_box_mask = false;
  // halo8: N = 8 structured box whose partner offsets are superincreasing in
  // the order x < y < z, so the stage kernels gather through the LDS halo
  // (gf::halo8_order's canonical order is GGFX's)
  bool halo8 = false;
  // reusable device workspace; ws_frozen: gf_mesh_reserve(GF_RESERVE_FREEZE)
  // was called, so gf_ws hands out only what is already allocated
  std::vector<void *> ws;
  std::vector<size_t> ws_cap;
  bool ws_frozen = false;
  // FP32 copies of the node metrics G[6] and the lumped mass for the
  // mixed-precision CG on embedded and deformed meshes (made at creation)
  float *G32[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  float *mass32 = nullptr;
  // members of a batched mesh (gf_mesh_set_segments): element ranges of
  // independent sub-meshes whose GGFX groups stay inside them; CG solves
  // give every member its own scalars and stopping test
  int nseg = 0;
  std::vector<int64_t> h_seg_eoff;  // nseg+1 element offsets (host)
  std::vector<int> h_seg_boff;      // nseg+1 update-block offsets (host)
  int64_t *seg_eoff = nullptr;      // device copies of the two
  int *seg_boff = nullptr;
  double *sscal = nullptr;          // per-member solver scalars (device)
  double *hsscal = nullptr, *hsscal_dev = nullptr;  // mapped host mirror of sscal
  std::vector<int> seg_iters;       // last solve: iterations per member
  std::vector<double> seg_rnorm, seg_rnorm0;
  // quadratic dealiasing (gf_dealias): the last order pq's interpolation
  // matrix and weights on the device
  int dealias_pq = -1;
  double *dl_I = nullptr, *dl_wq = nullptr;
  double *red = nullptr;    // reduction partials, red_cap doubles
  size_t red_cap = 0;
  double *scal = nullptr;   // solver scalars
  double *hscal = nullptr;  // pinned host mirror of scal
  double *hscal_dev = nullptr;  // device address of hscal (mapped)
  hipEvent_t cg_ev[2] = {nullptr, nullptr};  // CG batch completion (polling)
};

int gf_set_err(gf_ctx *ctx, int code, const char *msg);
int gf_check_hip(gf_ctx *ctx, hipError_t e, const char *what);

// Internal (C++ linkage) helpers shared across translation units.
int gf_dss_t(gf_mesh *m, int op, double *u);
// GGFX groups of a mesh of nbox boxes in sequence built on the device
// (gf_ggfx_dev.hip): fills m->goff, m->gmem, m->ngroups, m->maxgrp (and
// m->box for a single box).
int gf_box_groups_device(gf_mesh *m, int nbox, const int *ne, const int *periodic);
// Interpolation matrix p -> pq GLL nodes (column-major (pq+1) x (p+1)) and the
// order-pq weights, gf_setup.cpp.
int gf_dealias_matrix(int p, int pq, double *I, double *wq);
int gf_dss_f32(gf_mesh *m, int op, float *u);
int gf_helm_t(gf_mesh *m, double p, const double *pf, const double *qf, double qs,
              const double *u, double *uo);
int gf_grefderiv_t(gf_mesh *m, int idir, int trans, const double *u, double *du,
                   const double *edx);
int gf_mul_t(gf_mesh *m, const double *a, const double *b, double *out);
int gf_dot_t(gf_mesh *m, const double *a, const double *b, size_t n, double *result);
int gf_axpy_t(gf_mesh *m, double a, const double *x, double *y);
int gf_helm3c(gf_mesh *m, double p, const double *pf, const double *qf, double qs,
              const double *u, double *uo);

// Workspace: returns a device buffer of at least `bytes` for slot `slot`.
void *gf_ws(gf_mesh *m, int slot, size_t bytes);

#define GF_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return gf_check_hip(ctx, _e, #call); \
  } while (0)

#define GF_LAUNCH_CHECK(ctx) GF_HIP(ctx, hipGetLastError())

// Workspace slot numbers
enum {
  WS_TMP0 = 0, WS_TMP1, WS_TMP2, WS_TMP3, WS_TMP4, WS_TMP5,
  WS_RK_BEGIN = 8,          // RK: acc[3], K[3][3], u3[3]
  WS_CG_BEGIN = 32,         // CG vectors
  WS_ADJ_TMP = 48,          // adjoint RHS scratch (6 slots)
  WS_SLOTS = 64,            // kernel-timing scratch (16 slots)
  WS_ADJ_BEGIN = 96,        // adjoint stage states and cotangents (up to 36 slots)
  WS_ADJ_CKPT = 160         // adjoint checkpoints and segment states
};

// Reduction partials a mesh needs: one per apply block (at most one block
// per element) plus the update kernels' pairs.
size_t gf_red_need(size_t nelems);
