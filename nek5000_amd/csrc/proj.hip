// Residual projection around the Helmholtz solve (navier4.f:562-1325: hsolve,
// project1/project1_a, iproj_chk, proj_matvec, proj_ortho,
// proj_ortho_full_cgs2, project2/project2_a, givens_rotation).
//
// The space keeps up to mmx previous solutions X and their images B = A X in
// HBM (column k of X at xx + k*n), orthonormal in <x,b>_w = sum x*w*b.  Each
// Gram-Schmidt round is one pass that reads the columns once and writes
// partial sums over a fixed grid; the partials are summed in a fixed order on
// the device and the few scalars (Gram-Schmidt coefficients, norms, Givens
// pairs) come to the host, which mirrors the reference's gop reductions.  The
// column updates are single passes that apply the coefficients in the
// reference's order.  bbar = A xbar is not formed: the reference stores it
// but never reads it.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

namespace {

constexpr int kPjBlocks = 2048;
constexpr int kPjMax = 32;     // largest mmx (mxprev <= 68)
constexpr int kPjPairs = 8;    // dot products per pass of k_pj_dots
constexpr int kPjSlots = 2 * kPjMax + 2;

#define NK_LOOP(i, n)                                                          \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

struct PjPairs {
  const double *u[kPjPairs];
  const double *v[kPjPairs];
};

struct PjCoef {
  double c[kPjMax];
  int idx[kPjMax];
};

__device__ __forceinline__ void store_part(double a, double *part) {
  __shared__ double sa[4];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
  __syncthreads();
}

// partials of sum (u_j*w)*v_j (vlsc3) for the np pairs, slot j0+j
__global__ void __launch_bounds__(256)
    k_pj_dots(PjPairs pr, int np, const double *__restrict__ w, int64_t n, double *part,
              int j0) {
  double acc[kPjPairs];
#pragma unroll
  for (int j = 0; j < kPjPairs; ++j) acc[j] = 0.0;
  NK_LOOP(i, n) {
    const double wi = w[i];
#pragma unroll
    for (int j = 0; j < kPjPairs; ++j)
      if (j < np) acc[j] += (pr.u[j][i] * wi) * pr.v[j][i];
  }
#pragma unroll
  for (int j = 0; j < kPjPairs; ++j)
    if (j < np) store_part(acc[j], part + (int64_t)(j0 + j) * kPjBlocks);
}

// out[j] = fixed-order sum (op 0) or max (op 1) of the partials of slot j
__global__ void __launch_bounds__(256) k_pj_final(const double *part, double *out, int op) {
  __shared__ double sa[4];
  const double *p = part + (int64_t)blockIdx.x * kPjBlocks;
  double a = 0.0;
  for (int q = threadIdx.x; q < kPjBlocks; q += blockDim.x) a = op ? fmax(a, p[q]) : a + p[q];
  for (int o = 32; o > 0; o >>= 1) {
    const double b = __shfl_down(a, o, 64);
    a = op ? fmax(a, b) : a + b;
  }
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0)
    out[blockIdx.x] = op ? fmax(fmax(sa[0], sa[1]), fmax(sa[2], sa[3]))
                         : ((sa[0] + sa[1]) + sa[2]) + sa[3];
}

// y0 <- (init ? 0 : y0) + sum_j (s0*c_j) X0[idx_j];  y1 <- y1 + sum_j (s1*c_j)
// X1[idx_j], j in list order (add2s2 / cmult2 chains).  dot: 0 none, 1
// partials of (y0*w)*y1, 2 of (y1*w)*y1, after the update.
__global__ void __launch_bounds__(256)
    k_pj_comb(double *__restrict__ y0, const double *__restrict__ X0, double s0, int init,
              double *__restrict__ y1, const double *__restrict__ X1, double s1, PjCoef cf,
              int ns, const double *__restrict__ w, int dot, int64_t n, double *part) {
  double a = 0.0;
  NK_LOOP(i, n) {
    double v0 = init ? 0.0 : y0[i];
#pragma unroll
    for (int j = 0; j < kPjMax; ++j)
      if (j < ns) v0 = v0 + (s0 * cf.c[j]) * X0[(int64_t)cf.idx[j] * n + i];
    y0[i] = v0;
    if (y1) {
      double v1 = y1[i];
#pragma unroll
      for (int j = 0; j < kPjMax; ++j)
        if (j < ns) v1 = v1 + (s1 * cf.c[j]) * X1[(int64_t)cf.idx[j] * n + i];
      y1[i] = v1;
      if (dot == 1) a += (v0 * w[i]) * v1;
      if (dot == 2) a += (v1 * w[i]) * v1;
    }
  }
  if (dot) store_part(a, part);
}

// y0 *= c; y1 *= c (cmult)
__global__ void k_pj_scale2(double *__restrict__ y0, double *__restrict__ y1, double c,
                            int64_t n) {
  NK_LOOP(i, n) {
    y0[i] = y0[i] * c;
    y1[i] = y1[i] * c;
  }
}

// Column m-1 of X (and of Y, blockIdx.y = 1) scaled by sc, then the Givens
// sweep of proj_ortho: for k = m-1..1, rotate columns (k-1, k) by (c_k, s_k).
// Column k is final after its own rotation, so one carry register streams
// the sweep.
__global__ void __launch_bounds__(256)
    k_pj_givens(double *__restrict__ X, double *__restrict__ Y, int m, double sc, PjCoef cs,
                PjCoef sn, int64_t n) {
  double *__restrict__ A = blockIdx.y ? Y : X;
  NK_LOOP(i, n) {
    double carry = A[(int64_t)(m - 1) * n + i] * sc;
    for (int k = m - 1; k >= 1; --k) {
      const double c = cs.c[k], s = sn.c[k];
      const double xh = A[(int64_t)(k - 1) * n + i];
      const double t = c * xh + s * carry;
      A[(int64_t)k * n + i] = -s * xh + c * carry;
      carry = t;
    }
    A[i] = carry;
  }
}

// x <- x + xbar (when add); xm <- x
__global__ void k_pj_addcopy(double *__restrict__ x, const double *__restrict__ xbar, int add,
                             double *__restrict__ xm, int64_t n) {
  NK_LOOP(i, n) {
    const double v = add ? x[i] + xbar[i] : x[i];
    x[i] = v;
    xm[i] = v;
  }
}

// partial max of |h1 - h1old| and |h2 - h2old| (iproj_chk)
__global__ void __launch_bounds__(256)
    k_pj_maxdiff(const double *__restrict__ h1, const double *__restrict__ h1o,
                 const double *__restrict__ h2, const double *__restrict__ h2o, int64_t n,
                 double *part) {
  __shared__ double sa[4];
  double a = 0.0;
  NK_LOOP(i, n) a = fmax(a, fmax(fabs(h1[i] - h1o[i]), fabs(h2[i] - h2o[i])));
  for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_down(a, o, 64));
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sa[0], sa[1]), fmax(sa[2], sa[3]));
}

int pj_grid(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (int)(g < 8192 ? g : 8192);
}

// navier4.f:1308-1325 givens_rotation
void givens_rotation(double a, double b, double *c, double *s, double *r) {
  if (b != 0.0) {
    const double h = std::hypot(a, b);
    const double d = 1.0 / h;
    *c = std::fabs(a) * d;
    *s = std::copysign(d, a) * b;
    *r = std::copysign(1.0, a) * h;
  } else {
    *c = 1.0;
    *s = 0.0;
    *r = a;
  }
}

}  // namespace

struct Proj {
  int mesh = -1;
  int64_t n = 0;
  int mmx = 0, m = 0;
  double *h1old = nullptr, *h2old = nullptr, *xbar = nullptr;
  double *xx = nullptr, *bb = nullptr;
  double *part = nullptr;  // kPjSlots * kPjBlocks
  double *dout = nullptr;  // kPjSlots
  double *hout = nullptr;  // pinned, kPjSlots

  double *X(int k) const { return xx + (int64_t)k * n; }
  double *B(int k) const { return bb + (int64_t)k * n; }

  void release() {
    for (double *p : {h1old, h2old, xbar, xx, bb, part, dout}) (void)hipFree(p);
    if (hout) (void)hipHostFree(hout);
    h1old = h2old = xbar = xx = bb = part = dout = hout = nullptr;
  }

  // dots of the pairs, summed into hout[0..np)
  int dots(const std::vector<std::pair<const double *, const double *>> &pairs,
           const double *w, hipStream_t s) {
    const int np = (int)pairs.size();
    for (int j0 = 0; j0 < np; j0 += kPjPairs) {
      PjPairs pp{};
      const int cnt = np - j0 < kPjPairs ? np - j0 : kPjPairs;
      for (int j = 0; j < kPjPairs; ++j) {
        const auto &q = pairs[(size_t)(j0 + (j < cnt ? j : 0))];
        pp.u[j] = q.first;
        pp.v[j] = q.second;
      }
      hipLaunchKernelGGL(k_pj_dots, dim3(kPjBlocks), dim3(256), 0, s, pp, cnt, w, n, part, j0);
    }
    return fetch(np, 0, s);
  }

  int fetch(int np, int op, hipStream_t s) {
    hipLaunchKernelGGL(k_pj_final, dim3(np), dim3(256), 0, s, part, dout, op);
    NK_HIP(hipGetLastError());
    NK_HIP(hipMemcpyAsync(hout, dout, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    return NKH_OK;
  }

  // y0 (+)= sum s0 c_j X[idx_j]; y1 += sum s1 c_j B[idx_j]; optional dot into hout[0]
  int comb(double *y0, double s0, bool init, double *y1, double s1, const std::vector<int> &idx,
           const std::vector<double> &c, const double *w, int dot, hipStream_t s) {
    PjCoef cf{};
    const int ns = (int)idx.size();
    for (int j = 0; j < ns; ++j) {
      cf.c[j] = c[(size_t)j];
      cf.idx[j] = idx[(size_t)j];
    }
    hipLaunchKernelGGL(k_pj_comb, dim3(dot ? kPjBlocks : pj_grid(n)), dim3(256), 0, s, y0, xx,
                       s0, init ? 1 : 0, y1, bb, s1, cf, ns, w, dot, n, part);
    NK_HIP(hipGetLastError());
    if (dot) return fetch(1, 0, s);
    return NKH_OK;
  }
};

namespace {

std::mutex g_pj_mu;
std::vector<Proj *> g_projs;

Proj *get_proj(int id) {
  std::lock_guard<std::mutex> lk(g_pj_mu);
  if (id < 0 || id >= (int)g_projs.size()) return nullptr;
  return g_projs[(size_t)id];
}

// navier4.f:845-866 proj_matvec: b = mask * dssum(A x) with isd = 1
int proj_matvec(Mesh &m, double *b, const double *x, Coef c1, Coef c2, const double *msk,
                hipStream_t s) {
  NK_TRY(launch_axhelm(m, b, x, c1, c2, s));
  NK_TRY(gs_op(m.gs, b, 0, s));
  return mul_inplace(b, msk, m.n, s);
}

// navier4.f:1053-1121 proj_ortho_full_cgs2
int ortho_full_cgs2(Proj &P, const double *w, hipStream_t s) {
  const int m = P.m;
  if (m <= 0) return NKH_OK;
  const double tol = 1.e-7;
  std::vector<int> flag((size_t)m, 0);
  for (int it = 0; it < 2; ++it) {
    for (int k = m; k >= 1; --k) {
      std::vector<std::pair<const double *, const double *>> pr;
      for (int j = m; j >= k; --j) {
        pr.push_back({P.X(j - 1), P.B(k - 1)});
        pr.push_back({P.B(j - 1), P.X(k - 1)});
      }
      NK_TRY(P.dots(pr, w, s));
      std::vector<double> alpha((size_t)m + 1, 0.0);
      for (int j = m, q = 0; j >= k; --j, q += 2)
        alpha[(size_t)j] = .5 * (P.hout[q] + P.hout[q + 1]);
      std::vector<int> idx;
      std::vector<double> c;
      for (int j = m; j >= k + 1; --j) {
        idx.push_back(j - 1);
        c.push_back(alpha[(size_t)j]);
      }
      NK_TRY(P.comb(P.X(k - 1), -1.0, false, P.B(k - 1), -1.0, idx, c, w, 1, s));
      const double normp = std::sqrt(alpha[(size_t)k]);
      const double normk = std::sqrt(P.hout[0]);
      if (normk > tol * normp) {
        hipLaunchKernelGGL(k_pj_scale2, dim3(pj_grid(P.n)), dim3(256), 0, s, P.X(k - 1),
                           P.B(k - 1), 1.0 / normk, P.n);
        NK_HIP(hipGetLastError());
        flag[(size_t)k - 1] = 1;
      } else {
        flag[(size_t)k - 1] = 0;
      }
    }
  }
  int k = 0;
  for (int j = 1; j <= m; ++j) {
    if (flag[(size_t)j - 1] == 1) {
      k = k + 1;
      if (k < j) {
        NK_HIP(hipMemcpyAsync(P.X(k - 1), P.X(j - 1), sizeof(double) * P.n,
                              hipMemcpyDeviceToDevice, s));
        NK_HIP(hipMemcpyAsync(P.B(k - 1), P.B(j - 1), sizeof(double) * P.n,
                              hipMemcpyDeviceToDevice, s));
      }
    }
  }
  P.m = k;
  return NKH_OK;
}

// navier4.f:868-975 proj_ortho with the new column m
int ortho_new(Proj &P, const double *w, hipStream_t s) {
  const int m = P.m;
  if (m <= 0) return NKH_OK;
  std::vector<double> alpha((size_t)m + 1, 0.0), beta((size_t)m + 1, 0.0);
  std::vector<std::pair<const double *, const double *>> pr;
  for (int k = 1; k <= m; ++k) {
    pr.push_back({P.X(k - 1), P.B(m - 1)});
    pr.push_back({P.B(k - 1), P.X(m - 1)});
  }
  NK_TRY(P.dots(pr, w, s));
  for (int k = 1; k <= m; ++k)
    alpha[(size_t)k] = 0.5 * (P.hout[2 * (k - 1)] + P.hout[2 * (k - 1) + 1]);
  double nrm = std::sqrt(alpha[(size_t)m]);
  std::vector<int> idx;
  std::vector<double> c;
  for (int k = 1; k <= m - 1; ++k) {
    idx.push_back(k - 1);
    c.push_back(alpha[(size_t)k]);
  }
  NK_TRY(P.comb(P.X(m - 1), -1.0, false, P.B(m - 1), -1.0, idx, c, w, 0, s));
  pr.clear();
  for (int k = 1; k <= m - 1; ++k) {
    pr.push_back({P.X(k - 1), P.B(m - 1)});
    pr.push_back({P.B(k - 1), P.X(m - 1)});
  }
  if (!pr.empty()) NK_TRY(P.dots(pr, w, s));
  c.clear();
  for (int k = 1; k <= m - 1; ++k) {
    beta[(size_t)k] = 0.5 * (P.hout[2 * (k - 1)] + P.hout[2 * (k - 1) + 1]);
    c.push_back(beta[(size_t)k]);
  }
  // second round, then alpha(m) = (xx_m, bb_m)_w in the same pass
  NK_TRY(P.comb(P.X(m - 1), -1.0, false, P.B(m - 1), -1.0, idx, c, w, 1, s));
  for (int k = 1; k <= m - 1; ++k) alpha[(size_t)k] = alpha[(size_t)k] + beta[(size_t)k];
  alpha[(size_t)m] = std::sqrt(P.hout[0]);
  const double tol = 1.e-7;
  if (alpha[(size_t)m] > tol * nrm) {
    const double scl1 = 1.0 / alpha[(size_t)m];
    PjCoef cs{}, sn{};
    for (int k = m; k >= 2; --k) {
      const int h = k - 1;
      double cc, ss;
      givens_rotation(alpha[(size_t)h], alpha[(size_t)k], &cc, &ss, &nrm);
      alpha[(size_t)h] = nrm;
      cs.c[k - 1] = cc;
      sn.c[k - 1] = ss;
    }
    hipLaunchKernelGGL(k_pj_givens, dim3(pj_grid(P.n), 2), dim3(256), 0, s, P.xx, P.bb, m, scl1,
                       cs, sn, P.n);
    NK_HIP(hipGetLastError());
  } else {
    P.m = m - 1;
  }
  return NKH_OK;
}

int check_args(Proj *P, Mesh **mp, const void *v, const double *h1, const double *h2,
               const double *msk, const double *w, const char *who) {
  NK_CHECK(P, NKH_ERR_ARG, "%s: bad projection handle", who);
  Mesh *m = get_mesh(P->mesh);
  NK_CHECK(m && m->has_geom && m->has_gs && m->n == P->n, NKH_ERR_STATE,
           "%s: the mesh of the projection space is gone or changed", who);
  NK_CHECK(v && h1 && h2 && msk && w, NKH_ERR_ARG, "%s: null array", who);
  *mp = m;
  return NKH_OK;
}

}  // namespace

int proj_project1(int id, double *b, const double *h1, const double *h2, const double *msk,
                  const double *w, int ifmvbd, double *norms, int *ireset_out) {
  Proj *P = get_proj(id);
  Mesh *mp = nullptr;
  NK_TRY(check_args(P, &mp, b, h1, h2, msk, w, "project1"));
  Mesh &m = *mp;
  hipStream_t s = cur_stream();
  if (ireset_out) *ireset_out = 0;
  if (P->m <= 0) return NKH_OK;
  // iproj_chk (navier4.f:809-843)
  int ireset = 0;
  if (ifmvbd) {
    ireset = 1;
  } else {
    hipLaunchKernelGGL(k_pj_maxdiff, dim3(kPjBlocks), dim3(256), 0, s, h1, P->h1old, h2,
                       P->h2old, P->n, P->part);
    NK_TRY(P->fetch(1, 1, s));
    if (P->hout[0] > 0) {
      NK_HIP(hipMemcpyAsync(P->h1old, h1, sizeof(double) * P->n, hipMemcpyDeviceToDevice, s));
      NK_HIP(hipMemcpyAsync(P->h2old, h2, sizeof(double) * P->n, hipMemcpyDeviceToDevice, s));
      ireset = 1;
    }
  }
  if (ireset_out) *ireset_out = ireset;
  Coef c1, c2;
  if (ireset) {
    NK_TRY(classify_coefs(m, h1, h2, &c1, &c2, s));
    for (int j = 0; j < P->m; ++j) NK_TRY(proj_matvec(m, P->B(j), P->X(j), c1, c2, msk, s));
    NK_TRY(ortho_full_cgs2(*P, w, s));
  }
  // project1_a (navier4.f:737-807), bb4 = |b|_w with the first round's dots
  const int mm = P->m;
  std::vector<std::pair<const double *, const double *>> pr;
  for (int k = 0; k < mm; ++k) pr.push_back({P->X(k), b});
  pr.push_back({b, b});
  NK_TRY(P->dots(pr, w, s));
  const double bb4 = std::sqrt(P->hout[mm]);
  double baf = bb4;
  if (mm > 0) {
    std::vector<int> idx;
    std::vector<double> c;
    for (int k = 0; k < mm; ++k) {
      idx.push_back(k);
      c.push_back(P->hout[k]);
    }
    NK_TRY(P->comb(P->xbar, 1.0, true, b, -1.0, idx, c, w, 0, s));
    pr.pop_back();
    NK_TRY(P->dots(pr, w, s));
    for (int k = 0; k < mm; ++k) c[(size_t)k] = P->hout[k];
    NK_TRY(P->comb(P->xbar, 1.0, false, b, -1.0, idx, c, w, 2, s));
    baf = std::sqrt(P->hout[0]);
  }
  if (norms) {
    norms[0] = bb4;
    norms[1] = baf;
  }
  return NKH_OK;
}

int proj_project2(int id, double *x, const double *h1, const double *h2, const double *msk,
                  const double *w) {
  Proj *P = get_proj(id);
  Mesh *mp = nullptr;
  NK_TRY(check_args(P, &mp, x, h1, h2, msk, w, "project2"));
  Mesh &m = *mp;
  hipStream_t s = cur_stream();
  const int add = P->m > 0 ? 1 : 0;
  P->m = P->m + 1 < P->mmx ? P->m + 1 : P->mmx;
  hipLaunchKernelGGL(k_pj_addcopy, dim3(pj_grid(P->n)), dim3(256), 0, s, x, P->xbar, add,
                     P->X(P->m - 1), P->n);
  NK_HIP(hipGetLastError());
  Coef c1, c2;
  NK_TRY(classify_coefs(m, h1, h2, &c1, &c2, s));
  NK_TRY(proj_matvec(m, P->B(P->m - 1), P->X(P->m - 1), c1, c2, msk, s));
  return ortho_new(*P, w, s);
}

int proj_create(int mesh, int mxprev, int *id) {
  Mesh *m = get_mesh(mesh);
  NK_CHECK(id, NKH_ERR_ARG, "nkh_proj_create: null handle");
  NK_CHECK(m && m->has_geom && m->has_gs, NKH_ERR_STATE,
           "nkh_proj_create: mesh %d needs geometry and its gs map", mesh);
  const int mmx = (mxprev - 4) / 2;
  NK_CHECK(mmx >= 1 && mmx <= kPjMax, NKH_ERR_ARG,
           "nkh_proj_create: mxprev=%d gives %d vectors (1..%d supported)", mxprev, mmx,
           kPjMax);
  Proj *P = new Proj();
  P->mesh = mesh;
  P->n = m->n;
  P->mmx = mmx;
  auto alloc = [&](double **p, int64_t cnt) -> int {
    if (hipMalloc((void **)p, sizeof(double) * (size_t)cnt) != hipSuccess) {
      (void)hipGetLastError();
      set_error("nkh_proj_create: out of device memory (%lld doubles)", (long long)cnt);
      return NKH_ERR_NOMEM;
    }
    return NKH_OK;
  };
  int rc = NKH_OK;
  for (double **p : {&P->h1old, &P->h2old, &P->xbar}) rc = rc ? rc : alloc(p, P->n);
  rc = rc ? rc : alloc(&P->xx, (int64_t)mmx * P->n);
  rc = rc ? rc : alloc(&P->bb, (int64_t)mmx * P->n);
  rc = rc ? rc : alloc(&P->part, (int64_t)kPjSlots * kPjBlocks);
  rc = rc ? rc : alloc(&P->dout, kPjSlots);
  if (!rc && hipHostMalloc((void **)&P->hout, sizeof(double) * kPjSlots) != hipSuccess) {
    (void)hipGetLastError();
    set_error("nkh_proj_create: pinned host allocation failed");
    rc = NKH_ERR_NOMEM;
  }
  if (!rc) {
    // the reference's approx arrays start zeroed (h1old = h2old = 0)
    hipStream_t s = cur_stream();
    for (double *p : {P->h1old, P->h2old, P->xbar})
      if (hipMemsetAsync(p, 0, sizeof(double) * P->n, s) != hipSuccess) rc = NKH_ERR_HIP;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = NKH_ERR_HIP;
    if (rc) set_error("nkh_proj_create: clearing the space failed");
  }
  if (rc) {
    P->release();
    delete P;
    return rc;
  }
  std::lock_guard<std::mutex> lk(g_pj_mu);
  g_projs.push_back(P);
  *id = (int)g_projs.size() - 1;
  return NKH_OK;
}

int proj_destroy(int id) {
  std::lock_guard<std::mutex> lk(g_pj_mu);
  NK_CHECK(id >= 0 && id < (int)g_projs.size() && g_projs[(size_t)id], NKH_ERR_ARG,
           "nkh_proj_destroy: bad handle %d", id);
  Proj *P = g_projs[(size_t)id];
  (void)hipDeviceSynchronize();
  P->release();
  delete P;
  g_projs[(size_t)id] = nullptr;
  return NKH_OK;
}

int proj_info(int id, int *m, int *mmx) {
  Proj *P = get_proj(id);
  NK_CHECK(P, NKH_ERR_ARG, "nkh_proj_info: bad handle %d", id);
  if (m) *m = P->m;
  if (mmx) *mmx = P->mmx;
  return NKH_OK;
}

int proj_column(int id, int k, int which, const double **ptr) {
  Proj *P = get_proj(id);
  NK_CHECK(P && ptr, NKH_ERR_ARG, "nkh_proj_column: bad handle %d", id);
  NK_CHECK(which >= 0 && which <= 2, NKH_ERR_ARG, "nkh_proj_column: which=%d", which);
  NK_CHECK(which == 2 || (k >= 0 && k < P->mmx), NKH_ERR_ARG, "nkh_proj_column: k=%d", k);
  *ptr = which == 0 ? P->X(k) : which == 1 ? P->B(k) : P->xbar;
  return NKH_OK;
}

}  // namespace nekamd
