/*
 * nekamd.h -- C ABI of the MI355X implementation of Nek5000's
 * Helmholtz/Poisson solve path (axhelm, dssum/gs_op, cggo, hmholtz).
 *
 * Conventions
 *   - Every entry point returns NKH_OK (0) or a positive error code; the text
 *     of the last error is available from nkh_last_error().  Nek's Fortran
 *     shims turn a nonzero code into exitti(), keeping Nek's abort convention
 *     (comm_mpi.f:550-637).
 *   - Fields are Nek's (lx1,ly1,lz1,nel) column-major REAL*8 arrays: node
 *     (i,j,k) of element e at i + lx1*(j + lx1*(k + lx1*e)).
 *   - nkh_* entry points take HOST pointers (what Nek's Fortran passes by
 *     reference); nkd_* entry points take DEVICE pointers to arrays already
 *     resident in HBM.  Both run on the library's current HIP stream and
 *     return after the result is complete (nkh_*) or enqueued (nkd_*).
 *   - A "mesh" is the integer handle of one element set with its geometry and
 *     gather-scatter map: Nek's imsh=1 (nelv, velocity) or imsh=2 (nelt,
 *     temperature) meshes map to two handles.
 */
#ifndef NEKAMD_H
#define NEKAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NKH_OK 0
#define NKH_ERR_ARG 1         /* invalid argument or size */
#define NKH_ERR_HIP 2         /* HIP runtime error */
#define NKH_ERR_STATE 3       /* call out of order (no init, no geometry, ...) */
#define NKH_ERR_NOMEM 4       /* device allocation failed */
#define NKH_ERR_UNSUPPORTED 5 /* valid in Nek, not implemented here (2-D, axisymmetric) */

/* Solver knobs Nek reads from its commons inside hmholtz/cggo. */
typedef struct {
  int istep;       /* /TSTEP/ istep: cggo may stop at iteration 1 when istep<=5 (hmholtz.f:778) */
  int param18;     /* param(18): 1 -> plain residual norm, else B^-1-weighted (hmholtz.f:755-759) */
  double param22;  /* param(22): chktcg1 gate (hmholtz.f:54) and relative tol when <0 (:764) */
  double restol;   /* restol(ifield): overrides tin when nonzero and tin>0 (hmholtz.f:677) */
  int precision;   /* 0: FP64 (the drop-in); 1: multigrid smoother (fine-level local FDM
                      solves and overlap sum) in FP32 inside the FP64 Krylov iteration;
                      2, flexible CG: FP64 defect correction around single-precision
                      inner solves (FP32 operator, fine-level V-cycle and vectors; FP64
                      inner products, coarse solve and true residuals); 2, Jacobi PCG:
                      the iteration in FP32 (operator, vectors, preconditioner) with FP64
                      inner products, solution and residual refreshes (the exit test is
                      on the FP64 residual); both need uniform h1/h2, same tolerance */
  int param42;     /* param(42): PRES solver, 2 -> hmh_flex_cg with h1mg (hmholtz.f:646-653) */
  double param21;  /* param(21): PRES tolerance cap (>0) or relative tolerance (<0) (hmholtz.f:2206-2220) */
  int param40;     /* param(40): coarse solver of h1mg, 0 (XXT) -> solved to 1e-13, 1 (AMG) -> one
                      AMG V-cycle (navier8.f:209-217, crs_amg.c:129-) */
  /* residual projection switches read by hsolve (navier4.f:589-606) */
  int param93;     /* param(93): number of projection vectors; 0 turns projection off */
  int param94;     /* param(94): first step of projection for Helmholtz fields (0: off) */
  int param95;     /* param(95): first step of projection for the pressure (0: off) */
  int ifprojfld;   /* ifprojfld(ifield) of the field being solved */
  int ifield;      /* /TSTEP/ ifield; fields above ldimt_proj+1 are not projected */
  int ldimt_proj;  /* SIZE ldimt_proj */
  int ifmvbd;      /* moving mesh: the projection space is rebuilt every solve (navier4.f:818) */
} nkh_params;

/* What the reference leaves behind after a solve. */
typedef struct {
  int niter;      /* /iterhm/ niterhm */
  int converged;  /* rbn2 <= tol reached before maxit */
  double rbn0;    /* residual norm at iteration 1 */
  double rbn2;    /* residual norm at exit */
  double tol;     /* absolute tolerance used */
} nkh_result;

/* ---- lifetime ---------------------------------------------------------- */

/* Select the GPU and create the library stream.  Replaces nothing in the
   reference (Nek has no device); called once from the Fortran shim's setup. */
int nkh_init(int device);
int nkh_finalize(void);
const char *nkh_last_error(void);
/* Run subsequent work on the caller's stream (hipStream_t passed as void*);
   NULL selects HIP's null stream.  nkh_init starts on a library-owned
   stream, which nkh_use_library_stream() selects again. */
int nkh_set_stream(void *hip_stream);
int nkh_use_library_stream(void);
int nkh_synchronize(void);

/* ---- basis and numbering (host) --------------------------------------- */

/* speclib.f:107-122 ZWGLL: np GLL points and weights. */
int nkh_zwgll(double *z, double *w, int np);
/* speclib.f:800-833 DGLL: D(i,j) = phi_j'(z_i) and its transpose, ld = nz. */
int nkh_dgll(double *d, double *dt, const double *z, int nz);
/* navier8.f:2004-2360 SETVERT3D with ifcenter=.false. on one rank:
   vertex(2,2,2,nel) int64 -> glo_num(nx,nx,nx,nel) int64; *ngv = ids used. */
int nkh_setvert3d(int64_t *glo_num, int64_t *ngv, int nx, int nel, const int64_t *vertex);

/* ---- case files (host) --------------------------------------------------- */

/* reader_re2.f:543-620 READ_RE2_HDR: element counts and dimension of a .re2
   file (#v001..#v004, either byte order); *nfields = boundary-condition
   sections present. */
int nkh_re2_header(const char *path, int *nelgt, int *ndim, int *nelgv, int *nfields);
/* reader_re2.f:2-541 READ_RE2_DATA into Nek's /INPUT/ arrays: igroup(nel);
   xc, yc, zc(2^ndim, nel) in the file's (preprocessor) corner order (zc
   unused in 2-D); curve(5,12,nel) and ccurve(12,nel) (' ' where none);
   cbc(3 chars,6,nel,nfields) ('E  ' where no record) and bc(5,6,nel,nfields)
   for the first nfields boundary sections.  Any output may be NULL. */
int nkh_re2_read(const char *path, int nelgt, int ndim, int nfields, int *group, double *xc,
                 double *yc, double *zc, double *curve, char *ccurve, char *cbc, double *bc);
/* map2.f:712-942 READ_MAP, header: element count and active-element count
   of the case's .map (text) or .ma2 (binary, either byte order) file, the
   format chosen by the path's suffix. */
int nkh_map_header(const char *path, int *nel, int *nactive);
/* map2.f:712-942 READ_MAP: per element its partition (0-based, as genmap
   writes it) and its 2^ndim vertex ids, vertex(2^ndim, nel) in the
   hypercube corner order nkh_setvert3d reads (Nek's vertex array).  nel must
   equal the file's count.  Either output may be NULL. */
int nkh_map_read(const char *path, int nel, int ndim, int *part, int64_t *vertex);
/* Element vertex ids from corner coincidence (what Nek takes from the .ma2
   map, map2.f:712-942, or its parallel connectivity step): corners closer than
   tol_rel times the shortest element edge share an id, periodic faces
   (cbc 'P  ', bc(1..2) = connected element and face) join theirs.  3-D;
   vertex(2,2,2,nel) in the lexicographic corner order nkh_setvert3d reads;
   ids 1..*nvert in order of first appearance. */
int nkh_mesh_vertex_ids(int nel, int ndim, const double *xc, const double *yc, const double *zc,
                        const char *cbc, const double *bc, double tol_rel, int64_t *vertex,
                        int64_t *nvert);
/* genxyz.f:562-605 GENXYZ, 3-D: GLL coordinates xm1, ym1, zm1 (lx1^3 per
   element) from the corners and curve data of nkh_re2_read: LINQUAD
   (:1235-1400, xyzlin, and xyzquad with gh_face_extend for elements with
   'm' midside edges), then per element the 's' faces (SPHSRF :677-845;
   curve(1..4) = centre and radius) and the 'C' edges 1..8 (ARCSRF :2-103;
   curve(1) = signed radius).  'e' faces and other curve types return
   NKH_ERR_UNSUPPORTED. */
int nkh_mesh_xyz_from_corners(int lx1, int nel, const double *xc, const double *yc,
                              const double *zc, const char *ccurve, const double *curve,
                              double *xm1, double *ym1, double *zm1);

/* ---- meshes ------------------------------------------------------------ */

/* coef.f:1-359 GENWZ (3-D GLL part): creates a mesh of nel elements with
   lx1 points per direction and its 1-D basis. */
int nkh_mesh_create(int lx1, int nel, int *mesh);
int nkh_mesh_destroy(int mesh);
/* coef.f:879-927 XYZRST + :555-632 GLMAPM1 + :633-785 GEODAT1 on the GPU:
   g1m1..g6m1 and bm1 from the GLL coordinates (host arrays). */
int nkh_mesh_geom_from_xyz(int mesh, const double *xm1, const double *ym1, const double *zm1);
/* same, coordinates already on the device */
int nkd_mesh_geom_from_xyz(int mesh, const double *xm1, const double *ym1, const double *zm1);
/* Upload Nek's own /GMFACT/ g1m1..g6m1 and /MASS/ bm1 (host arrays). */
int nkh_mesh_set_geom(int mesh, const double *g1, const double *g2, const double *g3,
                      const double *g4, const double *g5, const double *g6, const double *bm1);
/* genxyz.f:351-470 SETDEF's IFDFRM for runs with param(59) = 0 (nel ints,
   nonzero = deformed).  The reference's axhelm and setprec skip the cross
   terms g4..g6 of an undeformed element (hmholtz.f:188-218: the fast path
   and `if (ifdfrm(e))`; :440-468), and sfastax (hmholtz.f:353-364) has
   overwritten its g4m1..g6m1 with g1/w..g3/w; the library then holds them
   at zero, now and after later geometry uploads.  NULL (what param(59) != 0
   and h1mg_setup's geom_reset give) treats every element as deformed again
   for later uploads. */
int nkh_mesh_set_deformed(int mesh, const int *ifdfrm);
/* dssum.f:1-31 SETUPDS (fgslib_gs_setup) from Nek's glo_num (host), then
   coef.f:1500-1543 SETINVM (binvm1 = 1/dssum(bm1)) and volvm1 = sum(bm1). */
int nkh_mesh_gs_setup(int mesh, const int64_t *glo_num);
/* n = lx1^3*nel, vol = volvm1, nseg/nsingle = shared groups / unshared ids */
int nkh_mesh_info(int mesh, int64_t *n, double *vol, int64_t *nseg, int64_t *nsingle,
                  int *max_mult);
/* Device copies of bm1 / binvm1 (device pointers owned by the library). */
int nkh_mesh_ptrs(int mesh, const double **bm1, const double **binvm1);

/* ---- operators on device arrays ---------------------------------------- */

/* hmholtz.f:72-259 AXHELM (3-D): au = h1*A u + h2*B u, local (unassembled).
   isd is accepted for signature parity; only the non-axisymmetric path exists. */
int nkd_axhelm(int mesh, double *au, const double *u, const double *h1, const double *h2,
               int isd);
/* Kernel choice for the AXHELM applies of order lx1 (no counterpart in the
   reference, whose axhelm has one form): variant 0 stages the element in LDS,
   1 stages one t-plane at a time (3 lx1^2 instead of 3 lx1^3 doubles of LDS);
   2 (plane) and 3 (element) pack more lx1 x lx1 thread slabs per block where
   one leaves lanes of its last wavefront idle (lx1 = 5-7, 9, 10, 12-14;
   elsewhere they run 1 and 0); 4 (element in LDS) and 5 (groups of t-planes
   in LDS) run the r- and s-direction contractions as FP64 matrix-core
   products (v_mfma_f64_16x16x4_f64, lx1 = 8-16; elsewhere they run 0).
   Variants 0-3 give the same plain apply bit for bit, as do 4 and 5, the two
   sets agreeing to rounding; the fused CG applies of 2 and 3 group their
   partial sums differently (iterates equal to rounding).  grid = blocks of
   the plain apply (0: one per element group).  Unless set, the first
   nkd_axhelm or solve of an order on a mesh of >= 2^20 nodes times the
   order's variants at two grids and keeps the fastest (*tuned = 1
   afterwards). */
int nkh_axhelm_tuning(int lx1, int *variant, int *grid, int *tuned);
/* sets the plain, CG-fused and flexible-CG-fused applies to variant */
int nkh_axhelm_set_tuning(int lx1, int variant, int grid);
/* the kernels of the CG-fused (cggo) and flexible-CG-fused applies, which
   the autotuner times separately on scratch vectors */
int nkh_axhelm_tuning_fused(int lx1, int *variant_cg, int *variant_flex);
/* Kernel of the single-precision operator of the mixed solvers
   (precision = 2) for order lx1 (no counterpart in the reference): 0 vector
   FMA, 1 the r- and s-direction contractions as FP32 matrix-core products
   (v_mfma_f32_16x16x4_f32, lx1 = 8-16, NKH_ERR_UNSUPPORTED elsewhere); the
   two agree to FP32 rounding.  The defaults are the faster of the two per
   order as measured on MI355X.  Process-wide. */
int nkh_axhelm32_set_variant(int lx1, int variant);
int nkh_axhelm32_variant(int lx1, int *variant);
/* Kernel of the local FDM solves (no counterpart in the reference, whose
   hsmg_do_fast has one form), for Schwarz blocks of up to 16 points: 0 vector
   FMA (default); for the single-precision smoother (precision 1 and 2) 1 the
   six contractions as FP32 matrix-core products (v_mfma_f32_16x16x4_f32), 2
   as bf16 matrix-core products with FP32 accumulation; for the FP64 smoother
   3 as FP64 matrix-core products (v_mfma_f64_16x16x4_f64).  Process-wide. */
int nkh_fdm_set_variant(int variant);
int nkh_fdm_variant(int *variant);
/* dssum.f:33-98 DSSUM: u_i <- sum over nodes sharing glo_num(i). */
int nkd_dssum(int mesh, double *u);
/* dssum.f:100-161 DSOP: op '+', '*', 'm' (min), 'M' (max). */
int nkd_dsop(int mesh, double *u, char op);
/* Adjoint coefficient gradient (no counterpart in the reference: the
   derivative of its axhelm form, hmholtz.f:188-225, with local_grad3,
   navier5.f:428-445).  Given the forward solution u of A(h1,h2) u = f and the
   adjoint lam of A lam = dJ/du (the same symmetric solve, e.g. nkd_hmholtz),
   writes per local node g1 = dJ/dh1 = -(D lam)^T G (D u) and, when g2 is not
   NULL, g2 = dJ/dh2 = -lam bm1 u.  For a continuous coefficient the gradient
   with respect to a shared node's value is nkd_dssum of these. */
int nkd_adjoint_coef_grad(int mesh, double *g1, double *g2, const double *u, const double *lam);
/* hmholtz.f:380-524 SETPREC: inverse of the assembled Jacobi diagonal. */
int nkd_setprec(int mesh, double *d, const double *h1, const double *h2);
/* hmholtz.f:2-69 HMHOLTZ: dssum(rhs); rhs*=mask; chktcg1; cggo.  name
   "PRES" takes cggo's branches to hmh_gmres (prm->param42 == 0, the
   reference's default) or hmh_flex_cg (param42 == 2) (hmholtz.f:639-653;
   both need nkh_mg_setup); param42 == 1 returns NKH_ERR_UNSUPPORTED.  rhs is modified in place as in
   the reference; u receives the solution.  history (optional) receives per
   Jacobi-PCG iteration (rbn2, rtz1, rho, alpha, beta), 5*min(maxit,900)
   doubles, or for PRES the residual norm per iteration, maxit doubles. */
int nkd_hmholtz(int mesh, const char *name, double *u, double *rhs, const double *h1,
                const double *h2, const double *mask, const double *mult, double tli,
                int maxit, int isd, const nkh_params *prm, nkh_result *res,
                double *history);
/* ---- H1 multigrid (PnPn pressure) ------------------------------------- */

/* hsmg.f:2234-2270 H1MG_SETUP (with hsmg.f:2272-2323 level orders, :2361-2394
   level numbering, :633-665 FDM, :3048-3104 Schwarz weights, fast3d.f:1542-1617
   swap_lengths and navier8.f:83-233 set_up_h1_crs) for a mesh that has its
   geometry and gs map.  Host arrays: fbc = 6 codes per element in symmetric
   face order (r-,r+,s-,s+,t-,t+) with get_fast_bc's meaning (fast3d.f:802-877:
   0 interior or periodic, 1 Dirichlet, 2 Neumann); vertex = the 8 corner ids
   per element (Nek's vertex(8,nel)); xm1/ym1/zm1 = GLL coordinates. */
int nkh_mg_setup(int mesh, const int *fbc, const int64_t *vertex, const double *xm1,
                 const double *ym1, const double *zm1);
/* hsmg.f:1855-1949 H1MG_SOLVE: z = M rhs.  rhs is masked in place (h1mg_mask
   at the fine level, as in the reference). */
int nkd_mg_solve(int mesh, double *z, double *rhs);
/* hmholtz.f:2164-2281 HMH_FLEX_CG: res holds the rhs on entry and the
   solution on return; wt = the multiplicity weight (vmult at the reference's
   call site hmholtz.f:653), pmask = the pressure mask.  Tolerance from
   chktcg1, param21 and istep as in the reference.  rnorm_history (optional,
   maxit doubles) receives the residual norm of every iteration. */
int nkd_hmh_flex_cg(int mesh, double *res, const double *h1, const double *h2,
                    const double *wt, const double *pmask, int maxit, const nkh_params *prm,
                    nkh_result *res_out, double *rnorm_history);
/* gmres.f:304-545 HMH_GMRES (right-preconditioned, restarted after 30 =
   lgmres steps): res holds the rhs on entry and the solution on return;
   tolerance and arguments as nkd_hmh_flex_cg.  rnorm_history (optional,
   maxit doubles) receives |gamma_{j+1}|/sqrt(vol) per Arnoldi step. */
int nkd_hmh_gmres(int mesh, double *res, const double *h1, const double *h2, const double *wt,
                  const double *pmask, int maxit, const nkh_params *prm, nkh_result *res_out,
                  double *rnorm_history);
/* ---- residual projection (hsolve) -------------------------------------- */

/* navier4.f:1179-1206 PROJ_GET_IVAR / hsolve's approx, napprox arrays: a
   projection space of mmx = (mxprev-4)/2 previous solutions and their images
   A x on the device, for fields of `mesh` (mxprev = SIZE mxprev, 20 by
   default; up to 68).  Starts empty with h1old = h2old = 0, as Nek's
   zero-initialised commons. */
int nkh_proj_create(int mesh, int mxprev, int *proj);
int nkh_proj_destroy(int proj);
/* napprox(2) (vectors held) and napprox(1) (= mmx) */
int nkh_proj_info(int proj, int *m, int *mmx);
/* Device pointer to column k of X (which 0), of B = A X (which 1), or to
   xbar (which 2, k ignored): read access for checks. */
int nkh_proj_column(int proj, int k, int which, const double **ptr);
/* navier4.f:636-735 PROJECT1 (+ iproj_chk :809, project1_a :737): b <- b
   minus its projection on the space, xbar kept for project2; rebuilds the
   space with the current h1/h2 when they changed (or ifmvbd).  msk = the
   field's mask, w = its vmult weight.  norms (optional, 2 doubles) receive
   the weighted norms of b before and after; *ireset (optional) 1 when the
   space was rebuilt.  Nothing happens while the space is empty. */
int nkd_project1(int proj, double *b, const double *h1, const double *h2, const double *msk,
                 const double *w, int ifmvbd, double *norms, int *ireset);
/* navier4.f:1123-1177 PROJECT2 (+ project2_a, proj_ortho :868-975): x <- x +
   xbar, then x joins the space (the oldest information leaves it once mmx
   vectors are held). */
int nkd_project2(int proj, double *x, const double *h1, const double *h2, const double *msk,
                 const double *w);
/* navier4.f:562-634 HSOLVE: hmholtz, or, when the switches in prm select
   projection (not ifprojfld / param93 / param94 / param95 / istep as in the
   reference), r*vmk, dssum, project1, hmhzpf (navier4.f:513-560: chktcg1
   then cggo with its PRES branches), project2.  proj < 0 means no space; it
   is an error when the switches select projection.  r is modified in place,
   u receives the solution.  res/history as nkd_hmholtz. */
int nkd_hsolve(int mesh, const char *name, double *u, double *r, const double *h1,
               const double *h2, const double *vmk, const double *vml, double tol, int maxit,
               int isd, int proj, const nkh_params *prm, nkh_result *res, double *history);

/* Hierarchy facts: number of levels, points per direction per level (nh[0]
   = level 1), coarse unknowns, mean coarse CG iterations per solve so far. */
int nkh_mg_info(int mesh, int *lmax, int *nh, int *ncrs, double *crs_avg_iters);

/* ---- batched solves --------------------------------------------------- */

/* Several hmholtz solves at once (no reference counterpart: Nek solves its
   velocity components and passive scalars one after another through
   hmholtz.f:2-69): a batch fixes the meshes of its solves (nb <= 64; a mesh
   may appear more than once, e.g. the three velocity components) and owns a
   work space, state and stream per solve. */
int nkh_batch_create(int nb, const int *meshes, int *batch);
int nkh_batch_destroy(int batch);
/* One solve of a batch: hmholtz's arguments (device arrays). */
typedef struct {
  double *u;           /* solution (out) */
  double *rhs;         /* right-hand side, dssum'd and masked in place */
  const double *h1, *h2, *mask, *mult;
  double tli;          /* tolerance, < 0 relative, as hmholtz */
} nkh_batch_solve;
/* hmholtz.f:2-69 + the Jacobi branch of cggo (:611-846) for every solve of
   the batch (non-PRES names): the set-up of each solve in order, then all
   solves iterate together -- eight steady-state iterations of the whole
   batch (each solve's launches on its own stream) are captured once in a
   hipGraph and replayed until every solve has stopped.  Iterates and
   results equal those of nkd_hmholtz on each solve (prm->precision 0 or
   2).  res: nb results; *graph_launches (optional): iterations enqueued
   for the slowest solve. */
int nkd_hmholtz_batch(int batch, const nkh_batch_solve *solves, int maxit,
                      const nkh_params *prm, nkh_result *res, int *graph_launches);

/* hmholtz.f:611-846 CGGO (Jacobi branch): f is the assembled, masked rhs
   (left unchanged), binv = binvm1/bintm1 of the mesh. */
int nkd_cggo(int mesh, double *x, const double *f, const double *h1, const double *h2,
             const double *mask, const double *mult, double tin, int maxit,
             const nkh_params *prm, nkh_result *res, double *history);

/* nkd_cggo without any device allocation: no counterpart in the reference.
   For callers that cannot allocate (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc).  The same FP64 Jacobi PCG
   (same kernels, bit-identical results) with the solve's vectors, partial
   sums, state and gs group flags in a static device array of the library
   (one per GPU, held until the process ends) instead of the mesh's lazily
   allocated work space; the array holds 6 vectors of 2^19 doubles plus
   partial sums, state and flags (about 26 MB).  Meshes of at most 2^19 =
   524,288 local nodes (e.g. 1,024 elements at lx1 = 8) only, else
   NKH_ERR_ARG; prm->precision must be 0.  Calls on one GPU share the array:
   calls that overlap on one GPU race. */
int nkd_cggo_noalloc(int mesh, double *x, const double *f, const double *h1, const double *h2,
                     const double *mask, const double *mult, double tin, int maxit,
                     const nkh_params *prm, nkh_result *res, double *history);

/* ---- timers ----------------------------------------------------------- */

/* Mirror of Nek's CTIMER counters (hmholtz.f:112-113 naxhm/taxhm, :26-31
   nhmhz/thmhz, dssum.f:52-95 tdsum).  When enabled, solves record HIP events
   around the axhelm and gather-scatter launches of each iteration that does
   work, on the stream those kernels run on.  name: "axhm", "dsum", "hmhz";
   in the flexible CG with h1mg also "fdm" (the fine-level local FDM solves
   of each V-cycle) and "mgv" (each V-cycle, on the fine level's stream). */
int nkh_timers_enable(int on);
int nkh_timers_reset(void);
int nkh_timers_read(const char *name, double *seconds, int64_t *calls);

/* ---- the same operators on host arrays (Fortran drop-in) ---------------- */

int nkh_axhelm(int mesh, double *au, const double *u, const double *h1, const double *h2,
               int isd);
int nkh_dssum(int mesh, double *u);
int nkh_hmholtz(int mesh, const char *name, double *u, double *rhs, const double *h1,
                const double *h2, const double *mask, const double *mult, double tli,
                int maxit, int isd, const nkh_params *prm, nkh_result *res,
                double *history);
int nkh_hsolve(int mesh, const char *name, double *u, double *r, const double *h1,
               const double *h2, const double *vmk, const double *vml, double tol, int maxit,
               int isd, int proj, const nkh_params *prm, nkh_result *res, double *history);

#ifdef __cplusplus
}
#endif

#endif /* NEKAMD_H */
