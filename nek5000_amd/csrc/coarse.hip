// Coarse-space solver: PCG with a smoothed-aggregation AMG V-cycle.
//
// Set-up (host): strength-of-connection graph, greedy aggregation (seed
// aggregates of a node and its free strong neighbours, leftovers joined to a
// neighbouring aggregate), tentative constant prolongator smoothed by one
// damped-Jacobi step, Galerkin product R A P; repeated until the problem is
// small enough for a dense inverse.
// Solve (device): one V(1,1)-cycle with damped Jacobi per CG iteration;
// kernels return at once when the CG state says it has converged, so a
// batch of iterations is enqueued ahead and the host reads the state once
// per batch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "coarse.h"

namespace nekamd {

namespace {

constexpr int kDenseMax = 600;   // last level at most this size
constexpr int kCrsBlocks = 1024; // max partials per reduction

struct Csr {
  int n = 0, m = 0;  // rows, cols
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
};

// ---------------------------------------------------------------------------
// host set-up
// ---------------------------------------------------------------------------
double diag_of(const Csr &A, int i) {
  for (int q = A.rp[i]; q < A.rp[i + 1]; ++q)
    if (A.ci[q] == i) return A.v[q];
  return 0.0;
}

// largest eigenvalue of D^-1 A by power iteration (a few steps, then 5%
// margin)
double lambda_max_dinv_a(const Csr &A, const std::vector<double> &d) {
  const int n = A.n;
  std::vector<double> x(n), y(n);
  for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.3 * std::sin(0.7 * i + 0.1);
  double lam = 1.0;
  for (int it = 0; it < 30; ++it) {
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) s += A.v[q] * x[A.ci[q]];
      y[i] = s / d[i];
      nrm += y[i] * y[i];
    }
    nrm = std::sqrt(nrm);
    double xn = 0.0;
    for (int i = 0; i < n; ++i) xn += x[i] * x[i];
    lam = nrm / std::sqrt(xn);
    for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
  }
  return 1.05 * lam;
}

// greedy aggregation on the strong graph; returns aggregate count
int aggregate(const Csr &A, double theta, std::vector<int> &agg) {
  const int n = A.n;
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = std::fabs(diag_of(A, i));
  auto strong = [&](int i, int q) {
    const int j = A.ci[q];
    return j != i && std::fabs(A.v[q]) >= theta * std::sqrt(d[i] * d[j]);
  };
  agg.assign(n, -1);
  int na = 0;
  // seed aggregates: a node with all strong neighbours free, plus them
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int q = A.rp[i]; q < A.rp[i + 1] && free_nb; ++q)
      if (strong(i, q) && agg[A.ci[q]] >= 0) free_nb = false;
    if (!free_nb) continue;
    agg[i] = na;
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q)
      if (strong(i, q)) agg[A.ci[q]] = na;
    ++na;
  }
  // leftovers join the aggregate of their strongest aggregated neighbour
  std::vector<int> snap = agg;
  for (int i = 0; i < n; ++i) {
    if (snap[i] >= 0) continue;
    double best = -1.0;
    int ba = -1;
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) {
      const int j = A.ci[q];
      if (j == i || snap[j] < 0 || !strong(i, q)) continue;
      const double sgt = std::fabs(A.v[q]);
      if (sgt > best) {
        best = sgt;
        ba = snap[j];
      }
    }
    agg[i] = ba;
  }
  // still alone: new aggregates with their free neighbours
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q)
      if (strong(i, q) && agg[A.ci[q]] < 0) agg[A.ci[q]] = na;
    ++na;
  }
  return na;
}

// P = (I - omega D^-1 A) P0, P0 the aggregate indicator
Csr smoothed_prolongator(const Csr &A, const std::vector<int> &agg, int na, double omega) {
  const int n = A.n;
  Csr P;
  P.n = n;
  P.m = na;
  P.rp.assign(n + 1, 0);
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    const double di = diag_of(A, i);
    row.clear();
    row.emplace_back(agg[i], 1.0);
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) row.emplace_back(agg[A.ci[q]], -omega * A.v[q] / di);
    std::sort(row.begin(), row.end(),
              [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                return a.first < b.first;
              });
    for (size_t q = 0; q < row.size();) {
      size_t t = q;
      double acc = 0.0;
      while (t < row.size() && row[t].first == row[q].first) acc += row[t++].second;
      if (acc != 0.0) {
        P.ci.push_back(row[q].first);
        P.v.push_back(acc);
      }
      q = t;
    }
    P.rp[i + 1] = (int32_t)P.ci.size();
  }
  return P;
}

Csr transpose(const Csr &P) {
  Csr R;
  R.n = P.m;
  R.m = P.n;
  R.rp.assign(P.m + 1, 0);
  for (int q = 0; q < (int)P.ci.size(); ++q) R.rp[P.ci[q] + 1]++;
  for (int i = 0; i < P.m; ++i) R.rp[i + 1] += R.rp[i];
  R.ci.resize(P.ci.size());
  R.v.resize(P.ci.size());
  std::vector<int32_t> pos(R.rp.begin(), R.rp.end() - 1);
  for (int i = 0; i < P.n; ++i)
    for (int q = P.rp[i]; q < P.rp[i + 1]; ++q) {
      const int J = P.ci[q];
      R.ci[pos[J]] = i;
      R.v[pos[J]] = P.v[q];
      pos[J]++;
    }
  return R;
}

// C = A * B (CSR, rows sorted)
Csr spgemm(const Csr &A, const Csr &B) {
  Csr C;
  C.n = A.n;
  C.m = B.m;
  C.rp.assign(A.n + 1, 0);
  std::vector<double> acc(B.m, 0.0);
  std::vector<int> mark(B.m, -1), cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) {
      const int k = A.ci[q];
      const double a = A.v[q];
      for (int t = B.rp[k]; t < B.rp[k + 1]; ++t) {
        const int j = B.ci[t];
        if (mark[j] != i) {
          mark[j] = i;
          acc[j] = 0.0;
          cols.push_back(j);
        }
        acc[j] += a * B.v[t];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols) {
      C.ci.push_back(j);
      C.v.push_back(acc[j]);
    }
    C.rp[i + 1] = (int32_t)C.ci.size();
  }
  return C;
}

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  if (hipMalloc(dst, sizeof(T) * std::max<size_t>(src.size(), 1)) != hipSuccess) {
    set_error("coarse solver allocation of %zu bytes failed", sizeof(T) * src.size());
    return NKH_ERR_NOMEM;
  }
  if (!src.empty())
    NK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return NKH_OK;
}

int dalloc(double **p, int64_t n) {
  if (hipMalloc(p, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) {
    set_error("coarse solver allocation of %lld doubles failed", (long long)n);
    return NKH_ERR_NOMEM;
  }
  return NKH_OK;
}

int to_ell(const Csr &A, EllDev &E) {
  int w = 1;
  for (int i = 0; i < A.n; ++i) w = std::max(w, A.rp[i + 1] - A.rp[i]);
  std::vector<int32_t> col((size_t)w * A.n);
  std::vector<double> val((size_t)w * A.n, 0.0);
  for (int i = 0; i < A.n; ++i) {
    const int len = A.rp[i + 1] - A.rp[i];
    for (int c = 0; c < w; ++c) {
      const size_t q = (size_t)c * A.n + i;
      if (c < len) {
        col[q] = A.ci[A.rp[i] + c];
        val[q] = A.v[A.rp[i] + c];
      } else {
        col[q] = 0;  // padding: value 0 at a valid column
      }
    }
  }
  E.n = A.n;
  E.width = w;
  NK_TRY(upload(&E.col, col));
  NK_TRY(upload(&E.val, val));
  return NKH_OK;
}

int to_csr(const Csr &A, CsrDev &C) {
  C.n = A.n;
  NK_TRY(upload(&C.rp, A.rp));
  NK_TRY(upload(&C.col, A.ci));
  NK_TRY(upload(&C.val, A.v));
  return NKH_OK;
}

void free_csr(CsrDev &C) {
  if (C.rp) (void)hipFree(C.rp);
  if (C.col) (void)hipFree(C.col);
  if (C.val) (void)hipFree(C.val);
  C = CsrDev();
}

void free_ell(EllDev &E) {
  if (E.col) (void)hipFree(E.col);
  if (E.val) (void)hipFree(E.val);
  E = EllDev();
}

// dense inverse of a small SPD matrix (Cholesky, then L^-T L^-1), col-major
bool dense_inverse(const Csr &A, std::vector<double> &inv) {
  const int n = A.n;
  std::vector<double> L((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int q = A.rp[i]; q < A.rp[i + 1]; ++q) L[(size_t)i + (size_t)n * A.ci[q]] = A.v[q];
  for (int j = 0; j < n; ++j) {
    double s = L[(size_t)j + (size_t)n * j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j + (size_t)n * k] * L[(size_t)j + (size_t)n * k];
    if (!(s > 0.0)) return false;
    const double d = std::sqrt(s);
    L[(size_t)j + (size_t)n * j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = L[(size_t)i + (size_t)n * j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i + (size_t)n * k] * L[(size_t)j + (size_t)n * k];
      L[(size_t)i + (size_t)n * j] = t / d;
    }
  }
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> e(n);
  for (int c = 0; c < n; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[c] = 1.0;
    for (int i = 0; i < n; ++i) {
      double s = e[i];
      for (int k = 0; k < i; ++k) s -= L[(size_t)i + (size_t)n * k] * e[k];
      e[i] = s / L[(size_t)i + (size_t)n * i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = e[i];
      for (int k = i + 1; k < n; ++k) s -= L[(size_t)k + (size_t)n * i] * e[k];
      e[i] = s / L[(size_t)i + (size_t)n * i];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i + (size_t)n * c] = e[i];
  }
  return true;
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
#define NK_ROWS(i, n) \
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

__device__ __forceinline__ void part_store(double a, double *part) {
  __shared__ double s1[4];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) s1[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((s1[0] + s1[1]) + s1[2]) + s1[3];
}

// fixed-order sum of nb partials by a block of 256 threads (valid on thread 0)
__device__ double part_sum(const double *part, int nb) {
  __shared__ double sa[4];
  double a = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) a += part[q];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  const double t = ((sa[0] + sa[1]) + sa[2]) + sa[3];
  __syncthreads();
  return t;
}

// x = wd*b; r = b - A (wd*b)
__global__ void __launch_bounds__(256)
    k_amg_pre(const int32_t *__restrict__ col, const double *__restrict__ val, int width, int n,
              const double *__restrict__ wd, const double *__restrict__ b, double *__restrict__ x,
              double *__restrict__ r, const CrsPcg *st) {
  if (st->done) return;
  NK_ROWS(i, n) {
    double s = 0.0;
    for (int c = 0; c < width; ++c) {
      const int64_t q = (int64_t)c * n + i;
      const int j = col[q];
      s += val[q] * (wd[j] * b[j]);
    }
    x[i] = wd[i] * b[i];
    r[i] = b[i] - s;
  }
}

// The PCG update of iteration k fused into the first smoothing step of the
// next V-cycle: alpha = r.z / p.w from the partial sums; x += alpha p;
// b = r_old - alpha w (formed on the fly for the neighbours too, with the
// same expression, and stored as the new residual of this row); then the
// smoothing step of k_amg_pre on b; partial sums of b.b for the exit test.
// r_old and r_new are different arrays (neighbours read r_old).
__global__ void __launch_bounds__(256)
    k_amg_pre_upd(const int32_t *__restrict__ col, const double *__restrict__ val, int width,
                  int n, const double *__restrict__ wd, const double *__restrict__ rold,
                  const double *__restrict__ w, const double *__restrict__ pn,
                  double *__restrict__ xs, double *__restrict__ rnew, double *__restrict__ x,
                  double *__restrict__ r, const CrsPcg *st, const double *prz,
                  const double *ppw, int nb, double *prr) {
  if (st->done) return;
  const double rz = part_sum(prz, nb);
  const double pAp = part_sum(ppw, nb);
  const double al = rz / pAp;
  double a = 0.0;
  NK_ROWS(i, n) {
    double s = 0.0;
    for (int c = 0; c < width; ++c) {
      const int64_t q = (int64_t)c * n + i;
      const int j = col[q];
      const double bj = rold[j] - al * w[j];
      s += val[q] * (wd[j] * bj);
    }
    xs[i] += al * pn[i];
    const double bi = rold[i] - al * w[i];
    rnew[i] = bi;
    x[i] = wd[i] * bi;
    r[i] = bi - s;
    a += bi * bi;
  }
  part_store(a, prr);
}

// y = x + wd*(b - A x); with part: partial sum of b.y
__global__ void __launch_bounds__(256)
    k_amg_post(const int32_t *__restrict__ col, const double *__restrict__ val, int width, int n,
               const double *__restrict__ wd, const double *__restrict__ b,
               const double *__restrict__ x, double *__restrict__ y, const CrsPcg *st,
               double *part) {
  if (st->done) return;
  double a = 0.0;
  NK_ROWS(i, n) {
    double s = 0.0;
    for (int c = 0; c < width; ++c) {
      const int64_t q = (int64_t)c * n + i;
      s += val[q] * x[col[q]];
    }
    const double yi = x[i] + wd[i] * (b[i] - s);
    y[i] = yi;
    a += b[i] * yi;
  }
  if (part) part_store(a, part);
}

// the same two smoothing steps with one wavefront per row of a CSR operator
// (levels whose rows are long and few)
__global__ void __launch_bounds__(256)
    k_amg_pre_w(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                const double *__restrict__ val, int n, const double *__restrict__ wd,
                const double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
                const CrsPcg *st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n;
       i += (gridDim.x * blockDim.x) >> 6) {
    double s = 0.0;
    for (int q = rp[i] + lane; q < rp[i + 1]; q += 64) {
      const int j = col[q];
      s += val[q] * (wd[j] * b[j]);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) {
      x[i] = wd[i] * b[i];
      r[i] = b[i] - s;
    }
  }
}

__global__ void __launch_bounds__(256)
    k_amg_post_w(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                 const double *__restrict__ val, int n, const double *__restrict__ wd,
                 const double *__restrict__ b, const double *__restrict__ x,
                 double *__restrict__ y, const CrsPcg *st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n;
       i += (gridDim.x * blockDim.x) >> 6) {
    double s = 0.0;
    for (int q = rp[i] + lane; q < rp[i + 1]; q += 64) s += val[q] * x[col[q]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) y[i] = x[i] + wd[i] * (b[i] - s);
  }
}

// v = M u (ADD: v += M u)
template <bool ADD>
__global__ void __launch_bounds__(256)
    k_ell_apply(const int32_t *__restrict__ col, const double *__restrict__ val, int width, int n,
                const double *__restrict__ u, double *__restrict__ v, const CrsPcg *st) {
  if (st->done) return;
  NK_ROWS(i, n) {
    double s = 0.0;
    for (int c = 0; c < width; ++c) {
      const int64_t q = (int64_t)c * n + i;
      s += val[q] * u[col[q]];
    }
    if constexpr (ADD) v[i] += s;
    else v[i] = s;
  }
}

// v = M u for a CSR matrix, one wavefront per row (lanes stride the row,
// fixed-order wave sum)
__global__ void __launch_bounds__(256)
    k_csr_apply(const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                const double *__restrict__ val, int n, const double *__restrict__ u,
                double *__restrict__ v, const CrsPcg *st) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < n;
       row += (gridDim.x * blockDim.x) >> 6) {
    double s = 0.0;
    for (int q = rp[row] + lane; q < rp[row + 1]; q += 64) s += val[q] * u[col[q]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) v[row] = s;
  }
}

// x = Ainv b with the symmetric dense inverse, one wavefront per row
__global__ void __launch_bounds__(256)
    k_dense_apply(const double *__restrict__ inv, int n, const double *__restrict__ b,
                  double *__restrict__ x, const CrsPcg *st) {
  if (st && st->done) return;
  const int lane = threadIdx.x & 63;
  for (int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < n;
       row += (gridDim.x * blockDim.x) >> 6) {
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += inv[(int64_t)n * row + j] * b[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) x[row] = s;
  }
}

// x = 0, r = b, partial r.r; state reset
__global__ void __launch_bounds__(256)
    k_pcg_init(double *x, double *r, double *p0, const double *b, int n, double *part) {
  double a = 0.0;
  NK_ROWS(i, n) {
    const double ri = b[i];
    x[i] = 0.0;
    r[i] = ri;
    p0[i] = 0.0;
    a += ri * ri;
  }
  part_store(a, part);
}

__global__ void __launch_bounds__(256)
    k_pcg_start(CrsPcg *st, const double *part_rr, int nb, double rtol, int maxit) {
  const double rr = part_sum(part_rr, nb);
  if (threadIdx.x == 0) {
    st->rr = rr;
    st->tol2 = rtol * rtol * rr;
    st->beta = 0.0;
    st->rz = 0.0;
    st->pb = 0;
    st->iter = 0;
    st->maxit = maxit;
    st->done = (rr == 0.0);
  }
}

__global__ void k_pcg_reset(CrsPcg *st) { st->done = 0; }

// Iteration k (1-based) of the coarse PCG.  Scalars are not carried in a
// mutable state: every block sums the partial sums it needs in the same
// fixed order (so all blocks agree bit for bit).  prz[k&1] holds the r.z
// partials of the V-cycle of iteration k (prz[0]: the first V-cycle);
// p of iteration k lives in cp[k&1], p of k-1 in cp[(k-1)&1].
//
// Exit test on the residual of iteration k-1 (by any block; block 0 records
// it for the host): rr <= rtol^2 rr0, or k-1 >= maxit.
__device__ __forceinline__ bool pcg_exit(CrsPcg *st, const double *prr, int nb, int k,
                                         double *rr_out) {
  if (k == 1) return st->done != 0;
  const double rr = part_sum(prr, nb);
  *rr_out = rr;
  return rr <= st->tol2 || k - 1 >= st->maxit;
}

// p = z + beta p_prev (own row, into cp[k&1]); w = A p with the neighbours'
// p formed on the fly; partial p.w
__global__ void __launch_bounds__(256)
    k_pcg_spmv(const int32_t *__restrict__ col, const double *__restrict__ val, int width, int n,
               double *__restrict__ w, double *cp0, double *cp1, const double *__restrict__ z,
               CrsPcg *st, const double *prz0, const double *prz1, const double *prr, int nb,
               int k, double *ppw) {
  double rr = 0.0;
  const bool ex = pcg_exit(st, prr, nb, k, &rr);
  if (ex) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && k > 1 && !st->done) {
      st->done = 1;
      st->iter = k - 1;
      st->rr = rr;
    }
    return;
  }
  double beta = 0.0;
  if (k > 1) {
    const double rz_new = part_sum((k - 1) & 1 ? prz1 : prz0, nb);
    const double rz_old = part_sum(k & 1 ? prz1 : prz0, nb);
    beta = rz_new / rz_old;
  }
  const double *po = (k - 1) & 1 ? cp1 : cp0;
  double *pn = k & 1 ? cp1 : cp0;
  double a = 0.0;
  NK_ROWS(i, n) {
    double s = 0.0;
    for (int c = 0; c < width; ++c) {
      const int64_t q = (int64_t)c * n + i;
      const int j = col[q];
      s += val[q] * (z[j] + beta * po[j]);
    }
    const double pi = z[i] + beta * po[i];
    pn[i] = pi;
    w[i] = s;
    a += pi * s;
  }
  part_store(a, ppw);
}

// the exit test alone (after a batch of iterations)
__global__ void __launch_bounds__(256) k_pcg_check(CrsPcg *st, const double *prr, int nb, int k) {
  double rr = 0.0;
  const bool ex = pcg_exit(st, prr, nb, k, &rr);
  if (ex && threadIdx.x == 0 && k > 1 && !st->done) {
    st->done = 1;
    st->iter = k - 1;
    st->rr = rr;
  }
}

inline int nblocks(int n) { return std::max(1, std::min((n + 255) / 256, kCrsBlocks)); }
// blocks for one wavefront per row
inline int wblocks(int n) { return std::max(1, std::min((n + 3) / 4, 8192)); }

}  // namespace

int CoarseSolver::setup(const std::vector<int32_t> &rowptr, const std::vector<int32_t> &col,
                        const std::vector<double> &val, hipStream_t s) {
  release();
  n = (int)rowptr.size() - 1;
  if (n <= 0) return NKH_OK;
  Csr A;
  A.n = A.m = n;
  A.rp = rowptr;
  A.ci = col;
  A.v = val;
  std::vector<Csr> As, Ps;
  As.push_back(A);
  double theta = 0.02;  // trilinear stencils: edge and corner couplings are weak but matter
  while (As.back().n > kDenseMax && (int)As.size() < 12) {
    const Csr &Ak = As.back();
    std::vector<double> d(Ak.n);
    for (int i = 0; i < Ak.n; ++i) d[i] = diag_of(Ak, i);
    std::vector<int> agg;
    const int na = aggregate(Ak, theta, agg);
    if (na >= Ak.n || na <= 0) break;  // no coarsening: stop here
    const double lmax = lambda_max_dinv_a(Ak, d);
    Csr P = smoothed_prolongator(Ak, agg, na, 4.0 / (3.0 * lmax));
    Csr R = transpose(P);
    Csr Ac = spgemm(R, spgemm(Ak, P));
    Ps.push_back(std::move(P));
    As.push_back(std::move(Ac));
    theta *= 0.5;
  }
  // the last level: dense inverse
  const Csr &Ad = As.back();
  std::vector<double> inv;
  NK_CHECK(Ad.n <= 4 * kDenseMax, NKH_ERR_STATE,
           "coarse solver: aggregation stalled at %d unknowns", Ad.n);
  NK_CHECK(dense_inverse(Ad, inv), NKH_ERR_STATE, "coarse solver: last level not SPD");
  ndense = Ad.n;
  {  // rows contiguous for the one-wavefront-per-row apply
    std::vector<double> rows(inv.size());
    for (int i = 0; i < ndense; ++i)
      for (int j = 0; j < ndense; ++j)
        rows[(size_t)ndense * i + j] = inv[(size_t)i + (size_t)ndense * j];
    NK_TRY(upload(&dinv_dense, rows));
  }
  NK_TRY(dalloc(&xd, ndense));
  NK_TRY(dalloc(&bd, ndense));
  nlevels_amg = (int)As.size() - 1;  // smoothed levels
  lev.resize(nlevels_amg);
  for (int k = 0; k < nlevels_amg; ++k) {
    AmgLevel &L = lev[k];
    const Csr &Ak = As[k];
    L.n = Ak.n;
    // rows long and few (Galerkin levels): one wavefront per row
    L.wide = k > 0 && Ak.ci.size() > (size_t)24 * Ak.n;
    if (L.wide) NK_TRY(to_csr(Ak, L.Aw));
    else NK_TRY(to_ell(Ak, L.A));
    NK_TRY(to_ell(Ps[k], L.P));
    NK_TRY(to_csr(transpose(Ps[k]), L.R));
    std::vector<double> d(Ak.n);
    for (int i = 0; i < Ak.n; ++i) d[i] = diag_of(Ak, i);
    const double lmax = lambda_max_dinv_a(Ak, d);
    std::vector<double> wd(Ak.n);
    for (int i = 0; i < Ak.n; ++i) wd[i] = (4.0 / (3.0 * lmax)) / d[i];
    NK_TRY(upload(&L.wd, wd));
    NK_TRY(dalloc(&L.x, L.n));
    NK_TRY(dalloc(&L.y, L.n));
    NK_TRY(dalloc(&L.r, L.n));
    NK_TRY(dalloc(&L.b, L.n));
  }
  if (nlevels_amg == 0) return NKH_OK;  // dense solve of the whole problem
  for (double **q : {&x, &rb[0], &rb[1], &z, &w, &p[0], &p[1]}) NK_TRY(dalloc(q, n));
  NK_TRY(dalloc(&part, 4 * kCrsBlocks));
  NK_HIP(hipMalloc(&st, sizeof(CrsPcg)));
  CrsPcg init{};
  NK_HIP(hipMemcpy(st, &init, sizeof(CrsPcg), hipMemcpyHostToDevice));
  (void)s;
  return NKH_OK;
}

void CoarseSolver::release() {
  for (AmgLevel &L : lev) {
    free_ell(L.A);
    free_csr(L.Aw);
    free_ell(L.P);
    free_csr(L.R);
    for (double *q : {L.wd, L.b, L.x, L.y, L.r})
      if (q) (void)hipFree(q);
  }
  lev.clear();
  for (double *q : {dinv_dense, xd, bd, x, rb[0], rb[1], z, w, p[0], p[1], part})
    if (q) (void)hipFree(q);
  if (st) (void)hipFree(st);
  dinv_dense = xd = bd = x = rb[0] = rb[1] = z = w = p[0] = p[1] = part = nullptr;
  st = nullptr;
  n = ndense = nlevels_amg = 0;
}

// z = M rhs: one V(1,1)-cycle; the partial sums of rhs.z land in part_out.
// With up, rhs is first formed as the PCG update r_old - alpha w of the
// level-0 smoothing kernel (k_amg_pre_upd).
void CoarseSolver::vcycle(const double *rhs, double *out, double *part_out, hipStream_t s,
                          const PcgUpdate *up) {
  const dim3 B(256);
  const double *bk = rhs;
  for (int k = 0; k < nlevels_amg; ++k) {
    AmgLevel &L = lev[k];
    if (k == 0 && up)
      hipLaunchKernelGGL(k_amg_pre_upd, dim3(nblocks(L.n)), B, 0, s, L.A.col, L.A.val,
                         L.A.width, L.n, L.wd, up->rold, up->w, up->pn, up->x, up->rnew, L.x,
                         L.r, st, up->prz, up->ppw, up->nb, up->prr);
    else if (L.wide)
      hipLaunchKernelGGL(k_amg_pre_w, dim3(wblocks(L.n)), B, 0, s, L.Aw.rp, L.Aw.col, L.Aw.val,
                         L.n, L.wd, bk, L.x, L.r, st);
    else
      hipLaunchKernelGGL(k_amg_pre, dim3(nblocks(L.n)), B, 0, s, L.A.col, L.A.val, L.A.width,
                         L.n, L.wd, bk, L.x, L.r, st);
    double *bnext = (k + 1 < nlevels_amg) ? lev[k + 1].b : bd;
    const int nnext = (k + 1 < nlevels_amg) ? lev[k + 1].n : ndense;
    hipLaunchKernelGGL(k_csr_apply, dim3(wblocks(nnext)), B, 0, s, L.R.rp, L.R.col, L.R.val,
                       nnext, L.r, bnext, st);
    bk = bnext;
  }
  hipLaunchKernelGGL(k_dense_apply, dim3(wblocks(ndense)), B, 0, s, dinv_dense, ndense, bd, xd,
                     st);
  const double *xc = xd;
  for (int k = nlevels_amg - 1; k >= 0; --k) {
    AmgLevel &L = lev[k];
    const int nbk = nblocks(L.n);
    hipLaunchKernelGGL(k_ell_apply<true>, dim3(nbk), B, 0, s, L.P.col, L.P.val, L.P.width, L.n,
                       xc, L.x, st);
    const double *bk2 = (k == 0) ? rhs : L.b;
    double *yk = (k == 0) ? out : L.y;
    if (L.wide)
      hipLaunchKernelGGL(k_amg_post_w, dim3(wblocks(L.n)), B, 0, s, L.Aw.rp, L.Aw.col,
                         L.Aw.val, L.n, L.wd, bk2, L.x, yk, st);
    else
      hipLaunchKernelGGL(k_amg_post, dim3(nbk), B, 0, s, L.A.col, L.A.val, L.A.width, L.n,
                         L.wd, bk2, L.x, yk, st, (k == 0) ? part_out : nullptr);
    xc = yk;
  }
}

// One V-cycle as the whole coarse solve (the reference's AMG coarse solver,
// param(40) = 1, applies its hierarchy once: crs_amg.c:129-).
int CoarseSolver::vcycle_only(double *xout, const double *b, hipStream_t s) {
  if (n == 0) return NKH_OK;
  if (nlevels_amg == 0) return solve(xout, b, s);
  hipLaunchKernelGGL(k_pcg_reset, dim3(1), dim3(1), 0, s, st);
  vcycle(b, xout, nullptr, s);
  NK_HIP(hipGetLastError());
  solves += 1;
  return NKH_OK;
}

int CoarseSolver::solve(double *xout, const double *b, hipStream_t s, double rtol_use) {
  if (n == 0) return NKH_OK;
  const dim3 B(256);
  if (nlevels_amg == 0) {  // small problem: the dense inverse is the solve
    hipLaunchKernelGGL(k_dense_apply, dim3(wblocks(n)), B, 0, s, dinv_dense, n, b, xout,
                       (const CrsPcg *)nullptr);
    NK_HIP(hipGetLastError());
    solves += 1;
    return NKH_OK;
  }
  const int nb0 = nblocks(n);
  double *p_rr = part, *p_pw = part + kCrsBlocks;
  double *p_rz0 = part + 2 * kCrsBlocks, *p_rz1 = part + 3 * kCrsBlocks;
  const EllDev &A0 = lev[0].A;
  // residual of iteration k in rb[k & 1]
  hipLaunchKernelGGL(k_pcg_init, dim3(nb0), B, 0, s, x, rb[0], p[0], b, n, p_rr);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), B, 0, s, st, p_rr, nb0,
                     rtol_use > 0.0 ? rtol_use : rtol, maxit);
  vcycle(rb[0], z, p_rz0, s);
  NK_HIP(hipGetLastError());
  CrsPcg hs{};
  int k = 1;
  int chunk = std::max(4, last_iters + 1);
  while (true) {
    for (int q = 0; q < chunk; ++q, ++k) {
      hipLaunchKernelGGL(k_pcg_spmv, dim3(nb0), B, 0, s, A0.col, A0.val, A0.width, n, w, p[0],
                         p[1], z, st, p_rz0, p_rz1, p_rr, nb0, k, p_pw);
      // the update (x += alpha p, r -= alpha w) runs inside the V-cycle's
      // first smoothing kernel
      PcgUpdate up{rb[(k - 1) & 1], w, (k & 1) ? p[1] : p[0], x, rb[k & 1],
                   ((k - 1) & 1) ? p_rz1 : p_rz0, p_pw, nb0, p_rr};
      vcycle(rb[k & 1], z, (k & 1) ? p_rz1 : p_rz0, s, &up);
    }
    hipLaunchKernelGGL(k_pcg_check, dim3(1), B, 0, s, st, p_rr, nb0, k);
    NK_HIP(hipGetLastError());
    NK_HIP(hipMemcpyAsync(&hs, st, sizeof(CrsPcg), hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
    NK_CHECK(k <= maxit + chunk + 1, NKH_ERR_STATE, "coarse PCG did not stop");
    chunk = 4;
  }
  last_iters = hs.iter;
  iters_total += hs.iter;
  solves += 1;
  NK_HIP(hipMemcpyAsync(xout, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  return NKH_OK;
}

}  // namespace nekamd
