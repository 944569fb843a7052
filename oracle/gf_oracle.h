/*
 * gf_oracle.h -- CPU restatement of GeoFLOW's spectral-element operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every routine restates the reference algorithm in the reference's own
 * operation order (left-to-right sums, std::accumulate from 0, serial dot
 * products) and is compiled with -ffp-contract=off so no FMA contraction
 * changes the rounding.  Each function cites the reference file:line it follows.
 *
 * Layout (same as the reference "local" vectors): element-contiguous,
 * field[e*Nx*Ny*Nz + k*Nx*Ny + j*Nx + i], x fastest
 * (src/cdg/sem/gmass.ipp:180-212, src/cdg/grid/ggrid_box.ipp:409-457).
 * Matrices are column-major: D[l + j*N] = dphi_j/dxi at node l
 * (src/cdg/blas/gtmatrix.hpp:104-107, src/cdg/sem/gllbasis.ipp:589-600).
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrGrid OrGrid;

/* ---- GLL basis (src/cdg/sem/gllbasis.ipp:410-747) ---------------------- */
/* xi[N], w[N], D[N*N] column-major, N = p+1.  Returns 0 on success. */
int or_gll_basis(int p, double *xi, double *w, double *D);

/* ---- Box grid (src/cdg/grid/ggrid_box.ipp:366-463, 753-978) ------------- */
/* bdy[2*dim]: 0 = GBDY_DIRICHLET ... 6 = GBDY_PERIODIC, 7 = GBDY_NONE
 * (src/cdg/include/gtypes.h:157), ordered x0,x1,y0,y1,z0,z1.
 * tol_fac: GGFX matching tolerance factor times minnodedist
 * (0.001 in src/apps/geoflow_cdg.cpp:772). */
OrGrid *or_grid_box_create(int dim, const int *p, const int *ne,
                           const double *P0, const double *P1,
                           const int *bdy, double tol_fac);
/* ---- Icosahedral 2-D sphere (src/cdg/grid/ggrid_icos.ipp) --------------- */
OrGrid *or_grid_icos_create(int p, int ilevel, double radius, double tol_fac);
/* refine: 0 GICOS_BISECTION, 1 GICOS_LAGRANGIAN */
OrGrid *or_grid_icos_create2(int p, int ilevel, int refine, double radius, double tol_fac);
/* ---- 2-D terrain-following GE_DEFORMED box (GGrid::add_terrain result,
 * GElem_base::dogeom2d/Jac/inv for GE_DEFORMED, gelem_base.cpp:877-955,
 * 1121-1136, 1249-1266); ref_inverse keeps the reference's transposed 2x2
 * inverse ------------------------------------------------------------------ */
OrGrid *or_grid_terrain_create(int p, const int *ne, const double *P0, const double *P1,
                               const int *bdy, double amp, int ref_inverse, double tol_fac);
/* new node coordinates for a 2-D deformed grid (no periodic directions):
 * geometry, minnodedist, GGFX, mass, volume recomputed (add_terrain,
 * ggrid.ipp:1522-1547) */
/* 3-D terrain-following GE_DEFORMED box (dogeom3d/Jac/inv,
 * gelem_base.cpp:972-1048, 1148-1168, 1268-1283) */
OrGrid *or_grid_terrain3_create(int p, const int *ne, const double *P0, const double *P1,
                                const int *bdy, double amp, double tol_fac);
int     or_grid_set_coords2d(OrGrid *g, const double *x, const double *y, int ref_inverse,
                             double tol_fac);
void    or_grid_destroy(OrGrid *g);

size_t or_grid_ndof(const OrGrid *g);
size_t or_grid_nelems(const OrGrid *g);
int    or_grid_dim(const OrGrid *g);
int    or_grid_gtype(const OrGrid *g);          /* 0 regular, 2 2d-embedded */
double or_grid_minnodedist(const OrGrid *g);
double or_grid_volume(const OrGrid *g);
/* Copy out grid arrays (caller-allocated).  which:
 *   "x","y","z"        node coordinates (ndof)
 *   "xp","yp","zp"     periodized node coordinates used for DSS matching (ndof)
 *   "mask"             Dirichlet mask (ndof)
 *   "mass"             lumped mass W x W x W * Jac (ndof)
 *   "imass"            inverse mass after DSS-Smooth (ndof)
 *   "jac"              Jacobian (ndof)
 *   "dxidx_ij"         dXi_i/dX_j for i,j in 0..2 (ndof)        */
int or_grid_get(const OrGrid *g, const char *which, double *out);
/* GGFX coincidence groups as CSR over nodes: for node n the group is
 * members[offs[n] .. offs[n+1]) in ascending order.  Pass NULL members
 * to get the total size in *nmem. */
int or_grid_ggfx(const OrGrid *g, size_t *offs, size_t *members, size_t *nmem);

/* ---- Tensor-product contractions (src/cdg/blas/gmtk.ipp:1404-2294) ------ */
/* dmxm: C(nai x nbj) = A(nai x naj) * B(nbi x nbj), column-major
 * (src/cdg/blas/cff_blas.F:846-943, dmxmp.F) */
void or_dmxm(double *C, const double *A, size_t nai, size_t naj,
             const double *B, size_t nbi, size_t nbj);
/* Reference-space derivative along idir (1..dim) of a single element-batched
 * field with per-direction sizes N[3]; D1,D2,D3 are the NxN derivative
 * matrices (NOT transposed); dotrans selects the transposed operator
 * (src/cdg/grid/ggrid.ipp:2213-2265). */
void or_grefderiv_raw(int dim, const int *N, size_t ne, const double *const *Dmats,
                      int idir, int dotrans, const double *u, double *du);

/* ---- Grid operators ----------------------------------------------------- */
void or_grefderiv(OrGrid *g, const double *u, int idir, int dotrans, double *du);
void or_grefdiv(OrGrid *g, const double *const *u, int dotrans, double *divu);
void or_deriv(OrGrid *g, const double *u, int idir, double *du);
double or_integrate(OrGrid *g, const double *u);
void or_mass_apply(OrGrid *g, const double *in, double *out);
void or_imass_apply(OrGrid *g, const double *in, double *out);
/* GGrid::dealias (src/cdg/grid/ggrid.ipp:2351-2503): quadratic dealiasing of
 * the product v1 v2 through the order-pq GLL grid (regular isotropic grids). */
void or_dealias(OrGrid *g, int pq, const double *v1, const double *v2, double *prod);
/* GHelmholtz::opVec_prod (src/cdg/sem/ghelmholtz.ipp:76-378):
 * p: Laplacian factor, np == 1 (scalar) or np == ndof (field);
 * q: mass factor field (length ndof) or NULL (pure Laplacian). */
void or_helmholtz_apply(OrGrid *g, const double *p, size_t np, const double *q,
                        const double *u, double *uo);
/* GGFX::doOp (src/cdg/comm/ggfx.hpp:344-443): op 0 Sum, 1 Smooth, 2 Max, 3 Min */
void or_dss(OrGrid *g, int op, double *u);
void or_imult(OrGrid *g, double *imult);
/* GAdvect::apply (src/cdg/sem/gadvect.ipp:57-99): nu velocity components */
void or_advect(OrGrid *g, const double *p, int nu, const double *const *u, double *po);
/* GMTK::constrain2sphere (2) (gmtk.ipp:2956-2980): in-place tangent projection of
 * (v0, v1, v2) with the reference's aliasing (the updated v0 feeds v1, both feed v2);
 * a no-op unless the grid is GE_2DEMBEDDED. */
void or_constrain2sphere(OrGrid *g, double *v0, double *v1, double *v2);
/* GBurgers::dudt_impl (src/cdg/pdes/gburgers.ipp:202-257).
 * mode 0 nonlinear (n_c = dim comps), 1 pure advection (u[0] advected by
 * u[1..nvel]), 2 heat. */
void or_burgers_rhs(OrGrid *g, int mode, double nu, int nstate, const double *const *u,
                    double *const *dudt);
/* GExRKStepper::step_b with Butcher RK of order `order` (1..4) driven by the
 * Burgers RHS and GGFX Smooth (src/cdg/pdes/gexrk_stepper.ipp:251-341,
 * src/cdg/pdes/gbutcherrk.ipp:80-127), then GBurgers::step_exrk's
 * constrain2sphere (src/cdg/pdes/gburgers.ipp:392-407).  For mode 1 the state
 * holds [u, c1, c2(, c3)] and only u evolves.  u updated in place. */
void or_burgers_step(OrGrid *g, int mode, double nu, int order, int nstate,
                     double t, double dt, double *const *u);

/* ---- GCG (src/cdg/solvers/gcg.ipp:239-429) ------------------------------ */
typedef struct {
  int    maxit;      /* LinSolverBase::Traits (src/pdeint/lin_solver_base.hpp:38-44) */
  double tol;
  int    normtype;   /* 0 INF, 1 EUC, 2 L2, 3 L1, 4 assembled 2-norm, pure relative */
  int    precond;    /* 0 identity (reference), 1 Jacobi (build-side addition)     */
} OrCGTraits;
/* Solves (p L + q M) x = b.  xb != NULL selects solve_impl(A,b,xb,x) with the
 * boundary mask.  resid (length >= maxit) receives the residual history.
 * Returns the GCGERR code (0 none, 1 NOCONVERGE, ...). */
int or_cg_solve(OrGrid *g, const double *p, size_t np, const double *q,
                const double *b, const double *xb, double *x,
                const OrCGTraits *tr, int *iters, double *resid);
/* Assembled Jacobi diagonal inverse for (p L + q M): 1/DSS_sum(diag). */
void or_helmholtz_diag_inv(OrGrid *g, const double *p, size_t np, const double *q,
                           double *dinv);

/* OpenMP thread count for the element loops (1 = serial reference order). */
void or_set_threads(int n);

/* Adjoint of the Burgers RHS and RK step (mode 0, regular grids): the
 * CPU reference for the hand-written GPU adjoint. */
void or_burgers_rhs_adjoint(OrGrid *g, double nu, int nstate, const double *const *u,
                            const double *const *w, double *const *lam);
void or_burgers_step_adjoint(OrGrid *g, double nu, int order, int nstate, double dt,
                             const double *const *un, double *const *lam);

#ifdef __cplusplus
}
#endif
#endif
