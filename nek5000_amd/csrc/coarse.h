// Coarse-space solver of the H1 multigrid: the assembled trilinear Galerkin
// problem (navier8.f:83-233 set_up_h1_crs) solved to a tight relative
// residual by conjugate gradients preconditioned with a smoothed-aggregation
// algebraic multigrid V-cycle (damped-Jacobi smoothing, dense inverse on the
// last level).  Stands in for the reference's exact XXT factor
// (crs_xxt.c:922-960): same solution to the CG tolerance.
#pragma once
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

// Sparse matrix in column-major ELL slots: entry c of row i at c*n + i.
struct EllDev {
  int n = 0, width = 0;
  int32_t *col = nullptr;
  double *val = nullptr;
};

// Sparse matrix in CSR (for rows too long and too few for ELL threads).
struct CsrDev {
  int n = 0;
  int32_t *rp = nullptr, *col = nullptr;
  double *val = nullptr;
};

struct AmgLevel {
  int n = 0;
  bool wide = false;     // operator applied from CSR, one wavefront per row
  EllDev A;              // level operator (thread per row)
  CsrDev Aw;             // level operator (wavefront per row), when wide
  EllDev P;              // prolongation to this level from the next (n x n_next)
  CsrDev R;              // restriction P^T (n_next x n), one wavefront per row
  double *wd = nullptr;  // omega / diag(A)
  double *b = nullptr, *x = nullptr, *y = nullptr, *r = nullptr;
};

// The PCG update folded into a V-cycle's first smoothing kernel: rnew =
// rold - alpha w, x += alpha pn, alpha = (prz sum) / (ppw sum); partial sums
// of r_new . r_new into prr.
struct PcgUpdate {
  const double *rold, *w, *pn;
  double *x, *rnew;  // rnew is the V-cycle's right-hand side
  const double *prz, *ppw;
  int nb;
  double *prr;
};

// Scalars of one coarse PCG solve, resident on the device.
struct CrsPcg {
  double rz, pAp, alpha, beta, rr, tol2;
  int iter, done, maxit, pb;
};

struct CoarseSolver {
  int n = 0;                 // unknowns
  std::vector<AmgLevel> lev; // lev[0] is the coarse problem itself
  int ndense = 0;            // size of the last level, solved by a dense inverse
  double *dinv_dense = nullptr;
  double *xd = nullptr, *bd = nullptr;
  // PCG vectors and state
  double *x = nullptr, *z = nullptr, *w = nullptr;
  double *rb[2] = {nullptr, nullptr};  // residual, alternating by iteration
  double *p[2] = {nullptr, nullptr};
  double *part = nullptr;  // 4 partial-sum arrays of kCrsBlocks
  CrsPcg *st = nullptr;
  double rtol = 1e-13;
  int maxit = 500;
  int last_iters = 0;
  int64_t iters_total = 0, solves = 0;
  int nlevels_amg = 0;

  // A: CSR of the assembled problem (rows sorted, diagonal present).
  int setup(const std::vector<int32_t> &rowptr, const std::vector<int32_t> &col,
            const std::vector<double> &val, hipStream_t s);
  // x = A^-1 b to the relative residual rtol (or rtol_use when > 0); host
  // waits on stream s.
  int solve(double *xout, const double *b, hipStream_t s, double rtol_use = 0.0);
  // x = M b, one AMG V-cycle (no host wait)
  int vcycle_only(double *xout, const double *b, hipStream_t s);
  void vcycle(const double *rhs, double *out, double *part_out, hipStream_t s,
              const PcgUpdate *up = nullptr);
  void release();
};

}  // namespace nekamd
