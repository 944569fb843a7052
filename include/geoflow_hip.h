/*
 * geoflow_hip.h -- C ABI of the MI355X-native GeoFLOW SEM operator path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No
 * exceptions cross this boundary and no torch types appear in it.  All field
 * pointers passed to operator entry points are DEVICE pointers holding
 * element-contiguous "local" vectors, exactly the layout of GeoFLOW's
 * GTVector state components:
 *     field[e*Nx*Ny*Nz + k*Nx*Ny + j*Nx + i]   (x fastest)
 * (reference: src/cdg/sem/gmass.ipp:180-212, src/cdg/grid/ggrid_box.ipp:409-457).
 * Matrices are column-major, D[l + j*N] = dphi_j/dxi at node l
 * (src/cdg/blas/gtmatrix.hpp:104-107, src/cdg/sem/gllbasis.ipp:589-600).
 *
 * Threading: one gf_ctx per host thread.  Operator calls are enqueued on the
 * context's HIP stream; they return once enqueued unless documented as
 * synchronous (those that return host scalars).  Hot-path entry points take
 * device scratch from the mesh's workspace, allocated on a mesh's first use
 * of that entry point or up front by gf_mesh_reserve; after that they do not
 * allocate.
 */
#ifndef GEOFLOW_HIP_H
#define GEOFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------
 * 0..4 mirror GCG::GCGERR (src/cdg/solvers/gcg.hpp:32). */
enum {
  GF_OK = 0,
  GF_ERR_NOCONVERGE = 1, /* GCGERR_NOCONVERGE */
  GF_ERR_SOLVE = 2,      /* GCGERR_SOLVE      */
  GF_ERR_PRECOND = 3,    /* GCGERR_PRECOND    */
  GF_ERR_BADITERNUM = 4, /* GCGERR_BADITERNUM */
  GF_ERR_ARG = 10,       /* invalid argument / shape mismatch */
  GF_ERR_HIP = 11,       /* HIP runtime error */
  GF_ERR_ALLOC = 12,     /* device allocation failed */
  GF_ERR_NONFINITE = 13, /* state not finite (GBurgers::step_impl assert, gburgers.ipp:327-331) */
  GF_ERR_UNSUPPORTED = 14
};

/* GBdyType (src/cdg/include/gtypes.h:157) */
enum {
  GF_BDY_DIRICHLET = 0, GF_BDY_INFLOW = 1, GF_BDY_NOSLIP = 2, GF_BDY_0FLUX = 3,
  GF_BDY_OUTFLOW = 4, GF_BDY_SPONGE = 5, GF_BDY_PERIODIC = 6, GF_BDY_NONE = 7
};

/* GElemType (src/cdg/include/gtypes.h:171) */
enum { GF_GE_REGULAR = 0, GF_GE_DEFORMED = 1, GF_GE_2DEMBEDDED = 2 };

/* GGFX reduction operators (src/cdg/comm/ggfx.hpp:36-76) */
enum { GF_DSS_SUM = 0, GF_DSS_SMOOTH = 1, GF_DSS_MAX = 2, GF_DSS_MIN = 3 };

/* LinSolverBase norm types (src/pdeint/lin_solver_base.hpp:38-44), plus
 * GF_NORM_ASSEMBLED_L2: 2-norm over unique (assembled) nodes with a purely
 * relative stopping rule ||r|| <= tol*||DSS(b)||. */
enum { GF_NORM_INF = 0, GF_NORM_EUC = 1, GF_NORM_L2 = 2, GF_NORM_L1 = 3, GF_NORM_ASSEMBLED_L2 = 4 };

typedef struct gf_ctx_s gf_ctx;
typedef struct gf_mesh_s gf_mesh;

/* ---- context ----------------------------------------------------------- */
/* stream: a hipStream_t (NULL = the legacy default stream). */
int  gf_ctx_create(int device, void *stream, gf_ctx **out);
int  gf_ctx_destroy(gf_ctx *ctx);
int  gf_ctx_set_stream(gf_ctx *ctx, void *stream);
int  gf_ctx_synchronize(gf_ctx *ctx);
/* Last error message for this context (static storage, never NULL). */
const char *gf_ctx_last_error(gf_ctx *ctx);
/* Library version string. */
const char *gf_version(void);

/* ---- host-side setup (no device work) ----------------------------------
 * GLLBasis<long double,double>::init (src/cdg/sem/gllbasis.ipp:714-747):
 * GLL nodes, weights and the derivative matrix D for order p (N = p+1). */
int gf_gll_basis(int p, double *xi, double *w, double *D);

/* GGridBox element and node generation for one rank
 * (src/cdg/grid/ggrid_box.ipp:254-463, 862-978), regular-element geometry
 * (GElem_base::dogeom2d/3d, src/cdg/sem/gelem_base.cpp:877-1048), boundary
 * mask (GGrid::init_bc_info, src/cdg/grid/ggrid.ipp:1410-1472) and periodized
 * coordinates (GGridBox::periodize, src/cdg/grid/ggrid_box.ipp:753-802).
 *   p[dim], ne[dim], P0[dim], P1[dim]; bdy[2*dim] ordered x0,x1,y0,y1(,z0,z1).
 * Outputs are caller-allocated (ndof = prod(ne) * prod(p+1)):
 *   xyz[dim]   node coordinates, each ndof (may be NULL)
 *   xyzp[dim]  periodized coordinates for DSS matching, each ndof (may be NULL)
 *   mask       ndof, 0 on Dirichlet boundary nodes (may be NULL)
 *   elem_dxidx nelems*dim, dXi_d/dX_d = 2/L_d per element (may be NULL)
 *   elem_jac   nelems, prod(L_d)/2^dim (may be NULL)
 *   minnodedist  GGrid::find_min_dist (src/cdg/grid/ggrid.ipp:977-1013) */
int gf_box_grid(int dim, const int *p, const int *ne, const double *P0, const double *P1,
                const int *bdy, double *const *xyz, double *const *xyzp, double *mask,
                double *elem_dxidx, double *elem_jac, double *minnodedist);

/* GGridIcos 2-D for one rank (src/cdg/grid/ggrid_icos.ipp:139-513): the
 * base icosahedron refined into 20 (nrows+1)^2 triangles (refine 0:
 * GICOS_BISECTION, nrows = 2^ilevel - 1; 1: GICOS_LAGRANGIAN, nrows = ilevel),
 * three GE_2DEMBEDDED quads per triangle, nodes projected to the sphere of
 * the given radius, and the embedded metrics of GElem_base::dogeom2d/Jac/inv
 * (src/cdg/sem/gelem_base.cpp:877-955, 1066-1170, 1249-1300).
 * Call with xyz == NULL to obtain *nelems; then ndof = nelems (p+1)^2 and
 *   xyz[3]    node coordinates
 *   dxidx[6]  dXi_i/dX_j at [i*3 + j], i < 2, j < 3 (entries may be NULL)
 *   jac       |det(dx/dxi)| / |x| per node (may be NULL)
 *   minnodedist  GGrid::find_min_dist (may be NULL) */
int gf_icos_grid(int p, int ilevel, int refine, double radius, size_t *nelems,
                 double *const *xyz, double *const *dxidx, double *jac, double *minnodedist);

/* GGrid::add_terrain's result for a 2-D box (src/cdg/grid/ggrid.ipp:1486-1548)
 * with an analytic terrain-following map in place of its CG solve, and the
 * GE_DEFORMED geometry of GElem_base::dogeom2d / Jac / inv
 * (src/cdg/sem/gelem_base.cpp:877-955, 1121-1136, 1249-1266): the regular box
 * of gf_box_grid (y not periodic) with every node moved to
 *   y' = y + h(x) (P1y - y)/(P1y - P0y),  h(x) = amp (1 - cos(2 pi (x-P0x)/Lx))/2.
 * dxidx[4] at [i*2 + j] = dXi_i/dX_j per node.  reference_inverse != 0 keeps
 * GElem_base::inv's 2x2 result, the transpose of dx/dxi^-1 (so the metrics
 * equal the reference's); 0 gives the true inverse.  jac = det(dx/dxi). */
int gf_terrain_box_grid(int p, const int *ne, const double *P0, const double *P1, const int *bdy,
                        double amp, int reference_inverse, double *const *xyz,
                        double *const *xyzp, double *mask, double *const *dxidx, double *jac,
                        double *minnodedist);

/* GE_DEFORMED geometry of 2-D elements from their node coordinates
 * (GElem_base::dogeom2d / Jac / inv, src/cdg/sem/gelem_base.cpp:877-955,
 * 1121-1136, 1249-1266), e.g. after GGrid::add_terrain has moved the nodes
 * (src/cdg/grid/ggrid.ipp:1528-1540 recomputes the geometry the same way).
 *   xyz[2]: nelems*(p+1)^2 coordinates; dxidx[4] at [i*2 + j] = dXi_i/dX_j;
 *   jac = det(dx/dxi); reference_inverse as gf_terrain_box_grid. */
int gf_deformed_geometry2d(int p, size_t nelems, const double *const *xyz, int reference_inverse,
                           double *const *dxidx, double *jac);

/* ---- GIO state files (src/cdg/io/gio.ipp) ---------------------------------
 * The layout GIO::read_header reads (gio.ipp:588-640) and GIO::sz_header
 * sizes (gio.ipp:678-690): int ivers, int dim, size_t nelems,
 * int porder[(ivers == 0 ? 1 : nelems) * dim], int gtype, size_t cycle,
 * double time, int multivar, size_t keys[nelems], then the field doubles
 * (one field, or every field in order when multivar != 0). */
typedef struct {
  int      ivers;     /* GIO version: 0 = one porder row, else one per element */
  int      dim;
  uint64_t nelems;
  int      gtype;     /* GF_GE_* */
  uint64_t cycle;
  double   time;
  int      multivar;  /* 0: one field per file */
  uint64_t nporder;   /* filled by gf_gio_read: ints in porder */
} gf_gio_info;
/* porder: (ivers == 0 ? 1 : nelems) * dim ints; keys: nelems ids or NULL
 * (0 .. nelems-1); fields: nfields host arrays of npts doubles. */
int gf_gio_write(const char *path, const gf_gio_info *info, const int *porder,
                 const uint64_t *keys, const double *const *fields, int nfields, size_t npts);
/* fields == NULL reads the header only; porder/keys (may be NULL) receive at
 * most porder_cap / keys_cap entries. */
int gf_gio_read(const char *path, gf_gio_info *info, int *porder, size_t porder_cap,
                uint64_t *keys, size_t keys_cap, double *const *fields, int nfields, size_t npts);

/* GE_DEFORMED geometry of 3-D elements from their node coordinates
 * (GElem_base::dogeom3d / Jac / inv, src/cdg/sem/gelem_base.cpp:972-1048,
 * 1148-1168, 1268-1283): dxidx[9] at [i*3 + j] = dXi_i/dX_j (the 3x3 inverse
 * is the true one), jac = |det(dx/dxi)|. */
int gf_deformed_geometry3d(int p, size_t nelems, const double *const *xyz,
                           double *const *dxidx, double *jac);
/* 3-D box of GE_DEFORMED hexahedra, bottom raised to
 * h(x,y) = amp (1 - cos(2 pi (x-P0x)/Lx)) (1 - cos(2 pi (y-P0y)/Ly)) / 4 by
 * z' = z + h (P1z - z)/(P1z - P0z); arguments as gf_terrain_box_grid in 3-D
 * (z not periodic), dxidx[9] at [i*3 + j]. */
int gf_terrain_box_grid3d(int p, const int *ne, const double *P0, const double *P1,
                          const int *bdy, double amp, double *const *xyz, double *const *xyzp,
                          double *mask, double *const *dxidx, double *jac, double *minnodedist);

/* GGFX::init local coincidence matching (src/cdg/comm/ggfx.hpp:167-342) on
 * host coordinates xyz[ncoord][ndof] with box tolerance tol.  Produces the
 * group of every node as CSR: members[offs[n] .. offs[n+1]) ascending.
 * Call with members == NULL to obtain *nmembers; offs needs ndof+1 entries.
 * Returns GF_ERR_ARG if a group exceeds max_dup (0 = unchecked). */
int gf_ggfx_build(size_t ndof, int ncoord, const double *const *xyz, double tol,
                  size_t max_dup, int64_t *offs, int64_t *members, size_t *nmembers);

/* GGFX::init local matching (src/cdg/comm/ggfx.hpp:167-342) for a 3-D box of
 * regular elements (elements x-fastest, N nodes per direction, ne[3]
 * elements, periodic[3] flags) on the device: the same CSR as gf_ggfx_build
 * on the box's coordinates, formed from the element lattice (no coordinate
 * matching).  offs (device, ndof+1 int64) and members (device int64, or NULL
 * to obtain *nmembers).  Every direction periodic or none: a box with both
 * kinds returns GF_ERR_UNSUPPORTED (periodize joins its groups differently).
 * Synchronous. */
int gf_box_ggfx_device(gf_ctx *ctx, int N, const int *ne, const int *periodic, int64_t *offs,
                       int64_t *members, size_t *nmembers);

/* ---- mesh (device-resident operator data) ------------------------------ */
typedef struct {
  int dim;               /* 2 or 3 (GDIM is a runtime value here)                  */
  int gtype;             /* GF_GE_REGULAR, GF_GE_2DEMBEDDED or GF_GE_DEFORMED (2-D, 3-D) */
  int N[3];              /* nodes per direction, p+1 (N[2] ignored when dim == 2)   */
  size_t nelems;
  const double *D[3];    /* host: column-major N x N derivative matrices            */
  const double *W[3];    /* host: GLL weights                                       */
  /* GE_REGULAR: per-element constants (host) */
  const double *elem_dxidx;  /* nelems*dim */
  const double *elem_jac;    /* nelems     */
  /* GE_2DEMBEDDED and GE_DEFORMED: per-node metrics (host), dXi_i/dX_j,
   * i < dim, j < nxy (3 embedded, dim deformed) at node_dxidx[i*3+j], and
   * node Jacobian (deformed: node_dxidx[i*3+j] for i, j < dim). */
  const double *node_dxidx[9];
  const double *node_jac;
  const double *xyz[3];  /* host node coordinates (optional; needed for sphere ops) */
  const double *mask;    /* host, ndof, or NULL (all ones)                          */
  /* GGFX groups from gf_ggfx_build (host) */
  const int64_t *ggfx_offs;
  const int64_t *ggfx_members;
  /* Optional box structure (3-D GE_REGULAR): elements per direction in
   * x-fastest order and per-direction periodicity, as given to gf_box_grid.
   * When set (box_ne[0] > 0) and the GGFX groups above are exactly the
   * groups this structure implies (checked at creation), fused kernels
   * assemble by gather instead of a separate DSS pass.  Zero = unstructured. */
  int box_ne[3];
  int box_periodic[3];
  /* Non-zero: build the GGFX groups on the device from the box structure
   * above (3-D GE_REGULAR isotropic box, every direction periodic or none)
   * instead of reading ggfx_offs/ggfx_members, which may then be NULL.  The
   * groups are those gf_ggfx_build forms from the box's coordinates
   * (gf_box_ggfx_device). */
  int box_device_maps;
  /* With box_device_maps: box_count >= 1 boxes in sequence (a batched mesh of
   * members, elements of box b following those of box b-1), box_list[6b ..
   * 6b+5] = ne0, ne1, ne2 and the three periodic flags of box b; the
   * groups never join two boxes.  box_ne/box_periodic are then unused and
   * the mesh has no single box structure. */
  int box_count;
  const int *box_list;
} gf_mesh_desc;

int gf_mesh_create(gf_ctx *ctx, const gf_mesh_desc *desc, gf_mesh **out);
int gf_mesh_destroy(gf_mesh *mesh);
size_t gf_mesh_ndof(const gf_mesh *mesh);
/* Mesh properties (diagnostic; no reference counterpart):
 * GF_MESH_BOX_GATHER  the structured-box gather assembly is in use;
 * GF_MESH_BOX_MASK    the boundary mask is formed from the box faces;
 * GF_MESH_CONST_D     3-D regular mesh whose D is the standard GLL matrix, so the
 *                     pencil kernels (D as compile-time constants) are used;
 * GF_MESH_BOX_HALO    p = 7 structured box whose RK stages gather the previous
 *                     stage derivative through an LDS halo of each element
 *                     (same results as the per-node gather);
 * GF_MESH_IMASS_ELEM  3-D regular mesh whose inverse mass is the same at every
 *                     element bit for bit (uniform periodic boxes), so the RK
 *                     stages read one element's values. */
enum { GF_MESH_BOX_GATHER = 1, GF_MESH_BOX_MASK = 2, GF_MESH_CONST_D = 4,
       GF_MESH_BOX_HALO = 8, GF_MESH_IMASS_ELEM = 16 };
int gf_mesh_flags(const gf_mesh *mesh);
/* Copy a device-resident mesh array to host: "mass", "imass", "imult",
 * "mask" (ndof doubles each).  Synchronous. */
int gf_mesh_get(gf_mesh *mesh, const char *which, double *host_out);

/* Workspace reservation (no reference counterpart: GeoFLOW's callers pass
 * scratch pools, State& utmp, to every operator; this library keeps them per
 * mesh).  Operators take device scratch from the mesh workspace, allocated
 * on first use.  gf_mesh_reserve allocates up front what the selected entry
 * points need, so later calls on this mesh allocate nothing:
 *   GF_RESERVE_RK       gf_burgers_step(s), gf_burgers_rhs, gf_advect, gf_deriv
 *   GF_RESERVE_CG       gf_cg_solve, FP64 (residual history of cg_maxit entries)
 *   GF_RESERVE_CG_MIXED gf_cg_solve, mixed precision (FP32 recurrence vectors)
 *   GF_RESERVE_ADJOINT  gf_burgers_step_adjoint, gf_burgers_adjoint with
 *                       adj_nsteps steps, checkpoints every adj_ckpt, nstate
 *                       components and RK order adj_order
 *   GF_RESERVE_FREEZE   from now on a call that would need more workspace
 *                       returns GF_ERR_ALLOC instead of allocating.
 * Synchronous. */
enum {
  GF_RESERVE_RK = 1,
  GF_RESERVE_CG = 2,
  GF_RESERVE_CG_MIXED = 4,
  GF_RESERVE_ADJOINT = 8,
  GF_RESERVE_FREEZE = 256
};
typedef struct {
  int nstate;      /* state components (RK, adjoint) */
  int cg_maxit;    /* CG iteration limit (residual history) */
  int adj_nsteps;  /* adjoint window length in steps */
  int adj_ckpt;    /* adjoint checkpoint interval */
  int adj_order;   /* adjoint RK order (1..4) */
} gf_reserve_params;
int gf_mesh_reserve(gf_mesh *mesh, int what, const gf_reserve_params *prm);

/* ---- device memory helpers (optional; any device allocation works) ----- */
int gf_malloc(gf_ctx *ctx, size_t bytes, void **dptr);
int gf_free(gf_ctx *ctx, void *dptr);
int gf_memcpy_h2d(gf_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */
int gf_memcpy_d2h(gf_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */

/* ---- tensor-product derivatives ----------------------------------------
 * GMTK::I3_X_I2_X_D1 / I3_X_D2_X_I1 / D3_X_I2_X_I1 (and 2-D forms) as
 * dispatched by GGrid::compute_grefderiv (src/cdg/grid/ggrid.ipp:2046-2265,
 * src/cdg/blas/gmtk.ipp:1404-2294).  idir in 1..dim; dotrans selects D^T. */
int gf_grefderiv(gf_mesh *mesh, int idir, int dotrans, const double *u, double *du);
/* GGrid::compute_grefderivs (src/cdg/grid/ggrid.ipp:1954-2016): du[dim]. */
int gf_grefderivs(gf_mesh *mesh, int dotrans, const double *u, double *const *du);
/* GGrid::compute_grefdiv (src/cdg/grid/ggrid.ipp:1847-1927): u[dim] (entries may be NULL). */
int gf_grefdiv(gf_mesh *mesh, int dotrans, const double *const *u, double *divu);
/* GGrid::deriv (src/cdg/grid/ggrid.ipp:1163-1191): physical derivative, idir 1..dim(+1 embedded). */
int gf_deriv(gf_mesh *mesh, int idir, const double *u, double *du);
/* GMTK::grad (src/cdg/blas/gmtk.ipp:2883-2892): du[k] = d u / d x_k for k < dim (+1 embedded). */
int gf_grad(gf_mesh *mesh, const double *u, double *const *du);

/* ---- mass -----------------------------------------------------------------
 * GMass::opVec_prod (src/cdg/sem/gmass.ipp:225-232); the inverse uses the
 * DSS-Smoothed mass (src/cdg/sem/gmass.ipp:75-82). */
int gf_mass_apply(gf_mesh *mesh, const double *in, double *out);
int gf_imass_apply(gf_mesh *mesh, const double *in, double *out);
/* GGrid::integrate (src/cdg/grid/ggrid.ipp:1077-1148).  Synchronous. */
int gf_integrate(gf_mesh *mesh, const double *u, double *result);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gf_integrate with its partial sums in a fixed 8 KiB
 * per-device scratch array (1024 doubles, held until the process ends)
 * instead of the mesh's reduction buffer.  result is a DEVICE pointer (one
 * double); asynchronous on the context stream.  Calls no hipMalloc/hipFree.
 * Any mesh size is accepted: the partial-sum kernel strides over the mesh
 * with at most 1024 blocks, one scratch slot each.  Every call on a device
 * shares that device's array, so calls that overlap on one device (two
 * streams, two host threads) race; calls ordered on one stream do not. */
int gf_integrate_noalloc(gf_mesh *mesh, const double *u, double *result);

/* Measured throughput peak (no reference counterpart; for the roofline
 * report): which = "mfma_f32" (v_mfma_f32_32x32x2f32), "mfma_f64"
 * (v_mfma_f64_16x16x4f64) or "valu_f64" (vector FMA), four waves per SIMD of
 * independent chains on every CU; *tflops receives the achieved TFLOP/s.
 * Synchronous; allocates a small output buffer. */
int gf_peak_probe(gf_ctx *ctx, const char *which, double *tflops);

/* GGrid::dealias (src/cdg/grid/ggrid.ipp:2351-2503): the quadratic
 * dealiasing of the product v1 v2 -- both interpolated to the order-pq GLL
 * grid of each element, multiplied there with the order-pq weights and
 * projected back (prod = Jac imass IQP^T W_q (IQP v1 . IQP v2)); used by
 * GDiv (src/cdg/sem/gdiv.ipp:91-128).  Regular meshes of constant p,
 * p <= pq <= 31; 2-D and 3-D.  The 3-D non-square tensor products are
 * those GMTK::D3_X_D2_X_D1 intends (its plane loop uses the output
 * plane stride, gmtk.ipp:1967-1968, so for pq > p the reference reads past
 * its first pass's result).  The first call for an order pq builds its
 * interpolation matrix on the device.  Asynchronous. */
int gf_dealias(gf_mesh *mesh, int pq, const double *v1, const double *v2, double *prod);

/* ---- Helmholtz ------------------------------------------------------------
 * GHelmholtz::opVec_prod -> reg_prod / embed_prod
 * (src/cdg/sem/ghelmholtz.ipp:76-378): uo = p L u + q M u, unassembled.
 * p: p_field (device, ndof) if non-NULL else the scalar p_scalar
 *    (set_Lap_scalar, ghelmholtz.ipp:590-603);
 * q: q_field (device, ndof) if non-NULL, else the scalar q_scalar; no mass
 *    term when q_field is NULL and q_scalar is 0 (set_mass_scalar,
 *    ghelmholtz.ipp:613-625; a one-entry q multiplies M u only when it is
 *    not 1, ghelmholtz.ipp:184-186, 370-373). */
int gf_helmholtz_apply(gf_mesh *mesh, double p_scalar, const double *p_field, double q_scalar,
                       const double *q_field, const double *u, double *uo);
/* Same operator in FP32 (inputs/outputs float) for the mixed-precision solver. */
int gf_helmholtz_apply_f32(gf_mesh *mesh, float p_scalar, float q_scalar, const float *q_field,
                           const float *u, float *uo);

/* ---- direct stiffness summation --------------------------------------------
 * GGFX::doOp (src/cdg/comm/ggfx.hpp:344-443), in place.  The sum of each
 * coincidence group runs over members in ascending local id from 0, as the
 * reference's std::accumulate does, so results are bit-identical to it. */
int gf_dss(gf_mesh *mesh, int op, double *u);
/* GGFX::get_imult (src/cdg/comm/ggfx.hpp:475-487): 1/multiplicity. */
int gf_imult(gf_mesh *mesh, double *imult);

/* ---- advection, Burgers RHS, RK step ---------------------------------------
 * GAdvect::apply (src/cdg/sem/gadvect.ipp:57-99): po = M sum_j u_j dp/dx_j. */
int gf_advect(gf_mesh *mesh, const double *p, int nvel, const double *const *u, double *po);
/* GBurgers::dudt_impl (src/cdg/pdes/gburgers.ipp:202-257).
 * mode 0: nonlinear, nstate = dim components (3 on sphere grids, where, as
 * in the reference's loop over GDIM, the third component's RHS is zero);
 * mode 1: pure advection of u[0] by u[1..nstate-1];  mode 2: heat. */
int gf_burgers_rhs(gf_mesh *mesh, int mode, double nu, int nstate, const double *const *u,
                   double *const *dudt);
/* GBurgers::step_impl -> GExRKStepper::step_b with Butcher RK of order
 * 1..4 and GGFX Smooth after each stage (src/cdg/pdes/gburgers.ipp:274-407,
 * src/cdg/pdes/gexrk_stepper.ipp:251-341, src/cdg/pdes/gbutcherrk.ipp:80-127).
 * u[nstate] device fields advanced in place.  check_finite != 0 performs the
 * reference's isfinite check and returns GF_ERR_NONFINITE on failure. */
int gf_burgers_step(gf_mesh *mesh, int mode, double nu, int order, int nstate, double t,
                    double dt, double *const *u, int check_finite);
/* nsteps consecutive steps (one host call; identical arithmetic). */
int gf_burgers_steps(gf_mesh *mesh, int mode, double nu, int order, int nstate, double t,
                     double dt, int nsteps, double *const *u);
/* Mixed-precision form of gf_burgers_steps for pure advection (mode 1,
 * nstate 4: u[0] advected by the velocity u[1..3]) on GE_2DEMBEDDED meshes:
 * the advected field, its RK stage fields, their DSS and the tensor-product
 * contractions in FP32; the metric terms (dXi/dX, G, mass, inverse mass and
 * the velocity-metric products) in FP64.  u[0] is FP64 at the boundary,
 * converted to FP32 at the start of the call and back at its end.  No
 * reference counterpart (GeoFLOW is FP64 throughout).  Other modes and
 * meshes return GF_ERR_UNSUPPORTED. */
int gf_burgers_steps_mixed(gf_mesh *mesh, int mode, double nu, int order, int nstate, double t,
                           double dt, int nsteps, double *const *u);

/* ---- adjoint (no reference counterpart: hand-written adjoint of the two
 * functions above, for gradients of final-time functionals) ---------------
 * Nonlinear mode, nstate = velocity components (dim, or 3 on embedded grids).
 * gf_burgers_rhs_adjoint: lam[k] += (dF/du)^T w at the state u, F the RHS of
 *   gf_burgers_rhs (GBurgers::dudt_impl, src/cdg/pdes/gburgers.ipp:203-257).
 * gf_burgers_step_adjoint: lam (in dJ/du_{n+1}) becomes dJ/du_n for one
 *   step of gf_burgers_step from un (step_b, gexrk_stepper.ipp:252-341);
 *   the stage states are recomputed.
 * gf_burgers_adjoint: for u(T) = nsteps steps from u0, lam (in dJ/du(T))
 *   becomes dJ/du(0); states are checkpointed every ckpt steps on the
 *   forward sweep and each segment is recomputed on the reverse sweep.  uT
 *   (optional) receives u(T).  Synchronous.  The checkpoints live in the
 *   mesh workspace ((ceil(nsteps/ckpt) + ckpt + 1) nstate fields), kept for
 *   later calls; gf_mesh_reserve(GF_RESERVE_ADJOINT) allocates them ahead. */
int gf_burgers_rhs_adjoint(gf_mesh *mesh, double nu, int nstate, const double *const *u,
                           const double *const *w, double *const *lam);
int gf_burgers_step_adjoint(gf_mesh *mesh, double nu, int order, int nstate, double dt,
                            const double *const *un, double *const *lam);
int gf_burgers_adjoint(gf_mesh *mesh, double nu, int order, int nstate, double dt, int nsteps,
                       int ckpt, const double *const *u0, double *const *lam,
                       double *const *uT);
/* gf_burgers_adjoint with the checkpoints persisted as GIO state files (the
 * GIO::write_state_impl layout of gf_gio_write, src/cdg/io/gio.ipp:588-690)
 * instead of device memory: checkpoint c, the state at step c*ckpt, is the
 * file "<prefix>.<c>.gio" (cycle = step, time = step*dt, one field per state
 * component); the reverse sweep reads them back.  Device memory holds one
 * segment ((ckpt + 1) nstate fields).  Results equal gf_burgers_adjoint's
 * bit for bit.  Synchronous.  No reference counterpart (the reference has
 * no adjoint); the files are GeoFLOW state files its tools can read. */
int gf_burgers_adjoint_gio(gf_mesh *mesh, double nu, int order, int nstate, double dt,
                           int nsteps, int ckpt, const double *const *u0, double *const *lam,
                           double *const *uT, const char *prefix);

/* ---- conjugate gradient -----------------------------------------------------
 * GCG::solve_impl (src/cdg/solvers/gcg.ipp:239-345, 361-382) applied to the
 * Helmholtz operator above. */
typedef struct {
  int    maxit;      /* default 512 */
  double tol;        /* default 1e-6 */
  int    normtype;   /* GF_NORM_* */
  int    precond;    /* 0: identity (reference);  1: Jacobi on the assembled diagonal */
  int    precision;  /* 0: FP64 CG;  1: FP32 recurrences with FP64 reliable updates */
  double inner_tol;  /* mixed: residual reduction between FP64 reliable updates (default 0.01) */
  double p_scalar;   /* Laplacian factor when p_field == NULL */
  const double *p_field;
  const double *q_field;
  double *resid;     /* host, optional, length >= maxit: residual history */
  double q_scalar;   /* mass factor when q_field == NULL (0: no mass term) */
} gf_cg_params;

typedef struct {
  int    status;     /* GF_OK or GF_ERR_NOCONVERGE ... */
  int    iters;      /* CG iterations (mixed: total inner iterations) */
  int    outer;      /* mixed: FP64 reliable updates */
  double rnorm;      /* final residual norm in the selected norm */
  double rnorm0;     /* norm of b used for the stopping rule */
} gf_cg_result;

/* b, xb, x: device fields.  xb != NULL selects solve_impl(A,b,xb,x): the
 * boundary mask is applied and xb is added back (gcg.ipp:361-382).
 * Synchronous (reads residual norms on the host each iteration batch).
 * Every option on regular, embedded and deformed meshes (the mixed
 * precision keeps FP32 copies of the node metrics in the mesh workspace). */
int gf_cg_solve(gf_mesh *mesh, const gf_cg_params *prm, const double *b, const double *xb,
                double *x, gf_cg_result *res);

/* Batched meshes (no reference counterpart): declare the mesh a union of
 * nseg independent members, member k being elements eoff[k] .. eoff[k+1]
 * (host int64, eoff[0] = 0, eoff[nseg] = nelems); no GGFX group may join two
 * members (checked).  gf_cg_solve then runs the GCG recurrence (or the
 * mixed-precision solver) with every member's own scalars, norm of b and
 * stopping test, all members in the same kernel launches; the result's
 * iters/rnorm/rnorm0 are the maxima over members, status GF_ERR_NOCONVERGE
 * if any member did not converge.  Supported on 3-D regular meshes with the
 * standard D and CSR assembly (the ensembles' concatenated meshes).  nseg 0
 * makes the mesh one system again. */
int gf_mesh_set_segments(gf_mesh *mesh, int nseg, const int64_t *eoff);
/* Per-member iterations and final/initial residual norms of the last
 * gf_cg_solve on a batched mesh (arrays of nseg, any may be NULL). */
int gf_cg_member_results(gf_mesh *mesh, int nseg, int *iters, double *rnorm, double *rnorm0);

/* Assembled Jacobi diagonal inverse of (p L + q M), q as in gf_helmholtz_apply;
 * on embedded and deformed meshes the diagonal includes the off-diagonal
 * metric terms.  No reference counterpart (GCG's default preconditioner is
 * the identity, gcg.ipp:38).  Returns GF_ERR_PRECOND (and leaves dinv
 * holding 1/diagonal) when an assembled diagonal entry is not positive, as
 * a sheared deformed element or p = q = 0 can make it. */
int gf_helmholtz_diag_inv(gf_mesh *mesh, double p_scalar, double q_scalar, const double *q_field,
                          double *dinv);

/* ---- diagnostics ------------------------------------------------------------
 * Average duration in ms of `reps` back-to-back launches of one kernel of the
 * path on scratch data of the mesh's size, timed with HIP events on the
 * context stream.  which: "burgers_stage" (fused RK stage, the stage 2-3
 * form), "cg_apply" / "cg_apply_f32" (CG direction update + Helmholtz apply),
 * "helmholtz".  No reference counterpart (the reference's timers are never
 * set, src/apps/hackathon_driver.cpp:91,258-259).  Synchronous. */
int gf_time_kernel(gf_mesh *mesh, const char *which, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* GEOFLOW_HIP_H */
