// Internal declarations shared by the translation units of libnekamd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/nekamd.h"

namespace nekamd {

// GLL points per direction the kernels are instantiated for: 2..kMaxNP.
constexpr int kMaxNP = 16;
// Fixed grid of the reduction kernels: partial sums have this many slots, so
// every reduction sums the same partials in the same order (deterministic).
constexpr int kRedBlocks = 2048;
// Partial-sum slots of the CG-fused axhelm (one per block of its fixed grid).
constexpr int kAxBlocks = 65536;
// Blocks of the CG assembly kernel (a multiple of the 8 XCDs).
constexpr int kGsBlocks = 8192;
constexpr int kMaxCgIter = 900;  // hmholtz.f:636 maxcg
// Blocks of the CG vector-update kernel (more blocks than the other
// reductions: the 6-read/2-write stream reaches its bandwidth with more
// waves in flight).
constexpr int kUpdBlocks = 16384;
// Layout of Mesh::red (doubles).
constexpr int kRedPart1 = 0;                         // (z,r) partials
constexpr int kRedPart2 = kUpdBlocks;                // (r,r)_binv partials
constexpr int kRedRho2 = 2 * kUpdBlocks;             // rho partials (unfused path)
constexpr int kRedScratch = kRedRho2 + kRedBlocks;   // setup reductions
constexpr int kRedOut = kRedScratch + kRedBlocks;    // 64 result slots
constexpr int kRedRho1 = kRedOut + 64;               // rho partials of axhelm
constexpr int kRedHist = kRedRho1 + kAxBlocks;       // 5*kMaxCgIter history
constexpr int kRedSize = kRedHist + 5 * kMaxCgIter;

void set_error(const char *fmt, ...);

#define NK_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::nekamd::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                          hipGetErrorString(e_));                                 \
      return NKH_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

#define NK_CHECK(cond, code, ...)                                                 \
  do {                                                                            \
    if (!(cond)) {                                                                \
      ::nekamd::set_error(__VA_ARGS__);                                           \
      return (code);                                                              \
    }                                                                             \
  } while (0)

#define NK_TRY(call)                                                              \
  do {                                                                            \
    int rc_ = (call);                                                             \
    if (rc_ != NKH_OK) return rc_;                                                \
  } while (0)

// Gather-scatter map over the local nodes of one mesh (gslib gs_op semantics
// on one rank).  Nodes with a nonzero global id form groups (members in
// increasing local index); groups are listed in the order of their first
// member.  Groups of one member (unshared ids) are listed too.
struct GsMap {
  int64_t n = 0;
  int64_t ngroups = 0;    // all nonzero ids
  int64_t nmem = 0;
  int64_t nseg = 0;       // groups with >= 2 members
  int64_t nsingle = 0;    // groups with one member
  int32_t *off = nullptr;     // device, ngroups+1
  int32_t *idx = nullptr;     // device, nmem
  uint8_t *flag = nullptr;    // device, ngroups (gs_prepare)
  bool interior_zero = false; // glo_num == 0 exactly at element-interior nodes
  int max_mult = 0;
  // packets: runs of whole groups of at most 64 members, one wavefront each
  // (groups [0, gpk)); groups [gpk, ngroups) have more than 64 members
  int64_t npk = 0, gpk = 0;
  int32_t *pk_m = nullptr;    // device, npk+1: first member of each packet
  int32_t *pk_g = nullptr;    // device, npk+1: first group of each packet
  uint64_t *pk_head = nullptr;  // device, npk: bit l set where a group starts at lane l
};

// State of one PCG solve, resident on the device.  The scalar kernels update
// it; the vector kernels read alpha/beta from it and return at once when
// `done` is set, so the host can enqueue iterations ahead without reading
// the residual back every iteration.
struct CgState {
  double rtz1, rtz2, rho, alpha, beta;
  double rbn0, rbn2, tol, tin, vol, param22;
  int iter;       // current iteration number (1-based)
  int maxit;      // min(maxit, 900)
  int done;
  int niter;
  int converged;
  int pending;    // x/r update of the previous iteration not yet applied
  int istep;
  int param18;
};

struct Mg;  // multigrid hierarchy (hsmg.hip)

struct Mesh {
  int np = 0;        // GLL points per direction (lx1)
  int64_t nel = 0;
  int64_t n = 0;     // np^3 * nel
  double D[kMaxNP * kMaxNP];   // D(i,j) column-major, as dxm1
  double w[kMaxNP];            // 1-D GLL weights
  double z[kMaxNP];            // 1-D GLL points
  double *Dd = nullptr;        // device copy of D
  double *g = nullptr;         // device, 6*n, factor-major: g[f*n + l]
  double *bm1 = nullptr;       // device
  double *binv = nullptr;      // device
  double vol = 0.0;            // sum of bm1
  bool has_geom = false;
  int geom_version = 0;        // bumped whenever g/bm1 change
  GsMap gs;
  bool has_gs = false;
  // solve workspace (allocated on first solve)
  double *ws = nullptr;
  size_t ws_doubles = 0;
  CgState *st = nullptr;       // device
  double *red = nullptr;       // device, partial sums
  Mg *mg = nullptr;            // H1 multigrid hierarchy (nkh_mg_setup)
  // FP32 copies of g and bm1 for the single-precision operator (built on
  // first use, rebuilt when geom_version moves)
  float *g32 = nullptr, *bm1_32 = nullptr;
  int g32_version = -1;
  // device, nel flags: 1 where the element is undeformed (IFDFRM false),
  // whose cross terms g4..g6 are then held at zero; null: all deformed
  uint8_t *undeformed = nullptr;
};

// CTIMER-style accumulators (seconds, calls), see nkh_timers_*.
struct Timers {
  bool on = false;
  double taxhm = 0, tdsum = 0, thmhz = 0;
  int64_t naxhm = 0, ndsum = 0, nhmhz = 0;
  // flexible CG + h1mg: the fine-level local FDM solves and the whole
  // V-cycle on the fine level's stream, per iteration
  double tfdm = 0, tmgv = 0;
  int64_t nfdm = 0, nmgv = 0;
};
Timers &timers();

// Elements per block of the element kernels (blocks of at least 64 lanes)
// and the fixed grid of the CG-fused axhelm (its number of rho partials).
inline int elems_per_block(int np) { return np * np >= 64 ? 1 : 64 / (np * np); }
// Sum of count partials by one block (blockDim.x a multiple of 64, at most
// 1024) in a fixed order: four interleaved accumulators per thread, so that
// several loads are in flight, then the waves in order.  Valid on every
// thread.  Used by the single-block scalar kernels of the Krylov loops.
__device__ inline double block_sum_partials(const double *__restrict__ part, int count) {
  __shared__ double sw[16];
  const int t = threadIdx.x, nt = blockDim.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int q = t;
  for (; q + 3 * nt < count; q += 4 * nt) {
    a0 += part[q];
    a1 += part[q + nt];
    a2 += part[q + 2 * nt];
    a3 += part[q + 3 * nt];
  }
  for (; q < count; q += nt) a0 += part[q];
  double v = (a0 + a1) + (a2 + a3);
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((t & 63) == 0) sw[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (nt >> 6); ++w) s += sw[w];
  __syncthreads();
  return s;
}

// A weight array read by the multigrid transfers and the final
// gather-scatter: FP64 or FP32 values, or one byte per node indexing a
// table of the level's distinct weights (lut).
__device__ __forceinline__ double wdec(const double *w, int64_t i, const double *) { return w[i]; }
__device__ __forceinline__ double wdec(const float *w, int64_t i, const double *) {
  return double(w[i]);
}
__device__ __forceinline__ double wdec(const uint8_t *w, int64_t i, const double *lut) {
  return lut[w[i]];
}
// a byte code decoded by arithmetic, without a table load: the weight 1/c
struct CountCode {
  uint8_t c;
};
__device__ __forceinline__ double wdec(const CountCode *w, int64_t i, const double *) {
  return 1.0 / (double)w[i].c;
}

inline int axhelm_cg_grid(const Mesh &m) {
  const int64_t g = (m.nel + elems_per_block(m.np) - 1) / elems_per_block(m.np);
  return (int)(g < kAxBlocks ? g : kAxBlocks);
}

Mesh *get_mesh(int id);
hipStream_t cur_stream();
int ensure_workspace(Mesh &m);  // solve vectors + state + partial sums
size_t cg_ws_doubles(int64_t n);  // work space of one Jacobi-PCG solve (doubles)
int ensure_small(Mesh &m);      // state + partial sums only
// Allocation-free Jacobi PCG: its buffers are a static device array sized for
// at most kNoAllocMaxN local nodes (cg.hip).
constexpr int64_t kNoAllocMaxN = (int64_t)1 << 19;
int noalloc_scratch(double **out);  // this device's address of the array
int run_cggo_noalloc(Mesh &m, double *x, const double *f, const double *h1, const double *h2,
                     const double *mask, const double *mult, double tin, int maxit,
                     const nkh_params &prm, nkh_result *res, double *history_host);

// ---- kernels.hip ----
// Coefficient description: field pointer or uniform scalar.
struct Coef {
  const double *f = nullptr;  // field (device) or nullptr
  double s = 0.0;             // value when uniform
  bool uniform() const { return f == nullptr; }
};

int launch_axhelm(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                  hipStream_t s);
// Kernel choice per order (0: element staged in LDS, 1: t-planes staged) and
// the plain apply's grid (0: one block per element group); tuned on the
// first apply of the order on a mesh of >= 2^20 nodes unless set first.
int axhelm_tuning(int np, int *variant, int *grid, int *tuned);
// Tune the order of m now if that is due (au is overwritten, u only read).
int axhelm_autotune(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                    hipStream_t s);
int axhelm_set_tuning(int np, int variant, int grid);  // all three applies
int axhelm_tuning_fused(int np, int *variant_cg, int *variant_flex);
// MODE 0 apply that returns while gate->done (replayed in captured loops).
int launch_axhelm_gated(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                        const CgState *gate, hipStream_t s);
// FP32 operator (mixed.hip): g32/bm1_32 current for the mesh's geometry.
int ensure_geom32(Mesh &m, hipStream_t s);
// FP32 Jacobi-CG apply: p = d*r + beta*p (beta = st->beta), w = A_L p with
// uniform h1/h2, FP64 partial sums of p.w per block; returns while st->done.
int launch_axhelm32_jac(const Mesh &m, float *w, float *p, const float *r, const float *d,
                        float h1, float h2, const CgState *st, double *rho_part,
                        hipStream_t s);
int launch_geom(Mesh &m, const double *x, const double *y, const double *z, hipStream_t s);
int launch_setprec(const Mesh &m, double *d, Coef h1, Coef h2, hipStream_t s);
// zero g4..g6 of the elements flagged undeformed (no-op without flags)
int apply_undeformed(Mesh &m, hipStream_t s);
// axhelm fused into the CG iteration: p = d*r + beta*p; w = A p (h1/h2 as
// above), rho partial over element-interior nodes when rho_part != nullptr.
int launch_axhelm_cg(const Mesh &m, double *w, double *p, const double *r, const double *d,
                     Coef h1, Coef h2, const double *mult, const CgState *st,
                     double *rho_part, hipStream_t s);

// axhelm fused into the flexible-CG iteration: p = beta*p + z (beta read
// from *betap), w = A p, rho partials of p . w over all local nodes; with x,
// also x += alpha p on the p read (alpha read from *alphap).
int launch_axhelm_flex(const Mesh &m, double *w, double *p, const double *z, Coef h1, Coef h2,
                       const double *betap, double *rho_part, hipStream_t s,
                       double *x = nullptr, const double *alphap = nullptr);

// ---- adjoint.hip ----
// g1 = -(D lam)^T G (D u), g2 = -lam*bm1*u per local node (g2 may be null)
int launch_adj_grad(const Mesh &m, double *g1, double *g2, const double *u, const double *lam,
                    hipStream_t s);

// ---- gs.hip ----
int gs_build(GsMap &gs, const int64_t *glo_num_host, int64_t n, int np, int64_t nel);
void gs_free(GsMap &gs);
// op: 0 sum, 1 product, 2 min, 3 max, 4 sum times 1/multiplicity (dsavg)
int gs_op(const GsMap &gs, double *u, int op, hipStream_t s);
// CG assembly: w <- mask * dssum(w) using the flags of gs_prepare; skipped
// when st is given and st->done.
// flag: group flags to use instead of gs.flag (a solve with its own mask)
// nblocks: grid (a multiple of 8; 0 -> kGsBlocks)
int gs_assemble(const GsMap &gs, double *w, const CgState *st, hipStream_t s,
                const uint8_t *flag = nullptr, int nblocks = 0);
int gs_assemble(const GsMap &gs, float *w, const CgState *st, hipStream_t s,
                const uint8_t *flag = nullptr);
// Per id group (single members included): S = sum of e; z_i <- mean over
// the members of S*w_i + z_i.  The fine-level end of the multigrid V-cycle.
int gs_fin(const GsMap &gs, double *z, const double *e, const double *w, hipStream_t s);
int gs_fin(const GsMap &gs, double *z, const float *e, const double *w, hipStream_t s);
int gs_fin(const GsMap &gs, float *z, const float *e, const float *w, hipStream_t s);
// the same with the weights as byte codes into lut
int gs_fin(const GsMap &gs, double *z, const double *e, const uint8_t *w, const double *lut,
           hipStream_t s);
int gs_fin(const GsMap &gs, float *z, const float *e, const uint8_t *w, const double *lut,
           hipStream_t s);
// Flags per group from mask/mult; *ok when every group has a 0/1 mask equal
// at all members and mult summing to 1 (red: kRedBlocks scratch).
int gs_prepare(GsMap &gs, const double *mask, const double *mult, double *red, bool *ok,
               hipStream_t s);

// ---- cg.hip ----
int run_cggo(Mesh &m, double *x, const double *f, Coef h1, Coef h2, const double *mask,
             const double *mult, double tin, int maxit, const nkh_params &prm,
             nkh_result *res, double *history_host);
// batched Jacobi-PCG solves (graph-captured iteration), see nkd_hmholtz_batch
int batch_create(int nb, const int *meshes, int *id);
int batch_destroy(int id);
int batch_meshes(int id, std::vector<int> *meshes);
int run_cggo_batch(int id, const double *const *x, const double *const *f, const Coef *h1,
                   const Coef *h2, const double *const *mask, const double *const *mult,
                   const double *tin, int maxit, const nkh_params &prm, nkh_result *res,
                   int *launches_out);
int run_chktcg1(Mesh &m, double *tol, const double *res, Coef h1, Coef h2,
                const double *mask, const double *mult);
int classify_coefs(Mesh &m, const double *h1, const double *h2, Coef *c1, Coef *c2,
                   hipStream_t s);
int run_setprec(Mesh &m, double *d, Coef h1, Coef h2, hipStream_t s);
int fused_conditions(Mesh &m, const double *mask, const double *mult, const double *f, bool *ok,
                     hipStream_t s);
int mul_inplace(double *a, const double *b, int64_t n, hipStream_t s);

// ---- hsmg.hip ----
// fbc: 6 codes per element in symmetric face order (0 interior/periodic,
// 1 Dirichlet, 2 Neumann); vertex: 8 ids per element; x/y/z: fine GLL
// coordinates.  All host arrays.
int mg_setup(Mesh &m, const int *fbc, const int64_t *vertex, const double *x, const double *y,
             const double *z);
void mg_free(Mesh &m);
// z = M rhs for rhs already masked with the fine multigrid mask
// crs_mode: 0 coarse problem solved to a 1e-13 relative residual, 1 one AMG
// V-cycle, 2 solved to 1e-4 (inner solves of the mixed solver)
// precision: 0 FP64; 1 the fine-level smoother (local FDM solves and overlap
// sum) in FP32, everything else FP64
int mg_solve(Mesh &m, double *z, const double *rhs, hipStream_t s, int crs_mode = 0,
             int precision = 0);
// the same V-cycle with the fine level in single precision throughout
// (input r and output z float; levels 2.. and the coarse solve FP64)
int mg_solve32(Mesh &m, float *z, const float *r, hipStream_t s, int crs_mode);
// Kernel of the single-precision local FDM solves (precision 1 and 2), for
// blocks of up to 16 points: 0 vector FMA (default), 1 FP32 matrix-core
// products, 2 bf16 matrix-core products with FP32 accumulation.
int fdm_set_variant(int v);
int fdm_variant();
int axhelm32_set_variant(int np, int v);
int axhelm32_variant(int np, int *v);
int mg_mask_fine(Mesh &m, double *r, hipStream_t s);
int mg_info(const Mesh &m, int *lmax, int *nh, int *ncrs, double *crs_avg_iters);
int run_flex_cg(Mesh &m, double *res, Coef h1, Coef h2, const double *wt, const double *mask,
                int maxit, const nkh_params &prm, nkh_result *out, double *rnorm_history);
// ---- mesh_io.cpp ----
int re2_header(const char *path, int *nelgt, int *ndim, int *nelgv, int *nfields);
int re2_read(const char *path, int nelgt, int ndim, int nfields, int *group, double *xc,
             double *yc, double *zc, double *curve, char *ccurve, char *cbc, double *bc);
int mesh_vertex_ids(int nel, int ndim, const double *xc, const double *yc, const double *zc,
                    const char *cbc, const double *bc, double tol_rel, int64_t *vertex,
                    int64_t *nvert);
int map_header(const char *path, int *nel, int *nactive);
int map_read(const char *path, int nel, int ndim, int *part, int64_t *vertex);
int mesh_xyz_from_corners(int lx1, int nel, const double *xc, const double *yc,
                          const double *zc, const char *ccurve, const double *curve,
                          double *xm1, double *ym1, double *zm1);

// ---- proj.hip ----
int proj_create(int mesh, int mxprev, int *id);
int proj_destroy(int id);
int proj_info(int id, int *m, int *mmx);
int proj_column(int id, int k, int which, const double **ptr);
int proj_project1(int id, double *b, const double *h1, const double *h2, const double *msk,
                  const double *w, int ifmvbd, double *norms, int *ireset_out);
int proj_project2(int id, double *x, const double *h1, const double *h2, const double *msk,
                  const double *w);
// ---- mixed.hip ----
// hmh_flex_cg to the same FP64 tolerance by defect correction: FP64 true
// residuals around inner flexible-CG solves in single precision.
int run_flex_cg_mixed(Mesh &m, double *res, Coef h1, Coef h2, const double *wt,
                      const double *mask, int maxit, const nkh_params &prm, nkh_result *out,
                      double *rnorm_history);
// ---- gmres.hip ----
int run_gmres(Mesh &m, double *res, Coef h1, Coef h2, const double *wt, const double *mask,
              int maxit, const nkh_params &prm, nkh_result *out, double *rnorm_history);

}  // namespace nekamd
