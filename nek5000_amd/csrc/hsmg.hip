// H1 multigrid preconditioner (hsmg.f h1mg_setup / h1mg_solve) and the
// flexible CG around it (hmholtz.f:2164-2281 hmh_flex_cg), on the GPU.
//
// Levels follow h1mg_setup_mg_nx (hsmg.f:2272-2323): orders 1, mglx2, lx1-1.
// Level-l fields are (nh,nh,nh,nel) element-major like the fine fields; the
// overlapping-Schwarz work arrays are (nh+2)^3 per element.  Every level has
// a gather-scatter map for its own numbering and one for the extended
// (nh+2)-point numbering (hsmg.f:2361-2394), built by setvert3d.
//
// Schwarz smoother at level l (hsmg.f:440-494):
//   k_sch_in   : W1 = ext(mask*r) with the ghost faces holding layer 2
//   gs(ext)    : W1 ghost += neighbours' ghost (sum of the two layers)
//   k_fdm      : ghost -= own layer 2 (= neighbour's layer 1), then the
//                fast-diagonalisation solve W2 = S D S^T W1 per element; the
//                pre-exchange ghost faces of W2 are kept in G
//   gs(ext)    : W2 ghost += neighbours' ghost
//   k_sch_out  : ghost -= G (= neighbour's solution at my layer n-2);
//                layer 2 += ghost; e = interior of W2
//   gs(level)  : e summed at shared nodes
//   k_sch_fin  : e *= mask * W_schwarz * sigma
// Coarse problem (navier8.f:83-233): trilinear Galerkin matrix assembled in
// CSR, solved by Jacobi-preconditioned CG to a tight relative tolerance.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

// ---------------------------------------------------------------------------
// host: 1-D operators
// ---------------------------------------------------------------------------
namespace mg1d {

// fast3d.f:1294-1349 FD_WEIGHTS_FULL (Fornberg), c(0:n,0:m) column-major
void fd_weights_full(double xx, const double *x, int n, int m, double *c) {
  auto C = [&](int j, int k) -> double & { return c[j + (n + 1) * k]; };
  double c1 = 1.0, c4 = x[0] - xx;
  for (int k = 0; k <= m; k++)
    for (int j = 0; j <= n; j++) C(j, k) = 0.0;
  C(0, 0) = 1.0;
  for (int i = 1; i <= n; i++) {
    const int mn = std::min(i, m);
    double c2 = 1.0, c5 = c4;
    c4 = x[i] - xx;
    for (int j = 0; j <= i - 1; j++) {
      const double c3 = x[i] - x[j];
      c2 = c2 * c3;
      for (int k = mn; k >= 1; k--) C(i, k) = c1 * (k * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
      C(i, 0) = -c1 * c5 * C(i - 1, 0) / c2;
      for (int k = mn; k >= 1; k--) C(j, k) = (c4 * C(j, k) - k * C(j, k - 1)) / c3;
      C(j, 0) = c4 * C(j, 0) / c3;
    }
    c1 = c2;
  }
}

// Generalised symmetric-definite eigenproblem A s = lam B s (hmholtz.f:1369
// generalev): Cholesky of B, cyclic Jacobi on L^-1 A L^-T, S = L^-T Q so that
// S^T B S = I; ascending eigenvalues.  a (n x n col-major) becomes S.
void generalev(double *a, const double *b, double *lam, int n) {
  double L[24 * 24], C[24 * 24], Q[24 * 24], T[24 * 24];
  auto M = [n](double *A, int i, int j) -> double & { return A[i + n * j]; };
  auto Mc = [n](const double *A, int i, int j) { return A[i + n * j]; };
  for (int j = 0; j < n; j++) {
    double s = Mc(b, j, j);
    for (int k = 0; k < j; k++) s -= M(L, j, k) * M(L, j, k);
    M(L, j, j) = std::sqrt(s);
    for (int i = j + 1; i < n; i++) {
      double t = Mc(b, i, j);
      for (int k = 0; k < j; k++) t -= M(L, i, k) * M(L, j, k);
      M(L, i, j) = t / M(L, j, j);
    }
    for (int i = 0; i < j; i++) M(L, i, j) = 0.0;
  }
  for (int c = 0; c < n; c++)
    for (int i = 0; i < n; i++) {
      double s = M(a, i, c);
      for (int k = 0; k < i; k++) s -= M(L, i, k) * M(T, k, c);
      M(T, i, c) = s / M(L, i, i);
    }
  for (int c = 0; c < n; c++)
    for (int i = 0; i < n; i++) {
      double s = M(T, c, i);
      for (int k = 0; k < i; k++) s -= M(L, i, k) * M(C, k, c);
      M(C, i, c) = s / M(L, i, i);
    }
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      const double s = 0.5 * (M(C, i, j) + M(C, j, i));
      M(C, i, j) = s;
      M(C, j, i) = s;
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) M(Q, i, j) = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += M(C, i, j) * M(C, i, j);
        if (i != j) off += M(C, i, j) * M(C, i, j);
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = M(C, p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (M(C, q, q) - M(C, p, p)) / (2.0 * apq);
        const double t =
            (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; k++) {
          const double x = M(C, k, p), y = M(C, k, q);
          M(C, k, p) = cs * x - sn * y;
          M(C, k, q) = sn * x + cs * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = M(C, p, k), y = M(C, q, k);
          M(C, p, k) = cs * x - sn * y;
          M(C, q, k) = sn * x + cs * y;
        }
        for (int k = 0; k < n; k++) {
          const double x = M(Q, k, p), y = M(Q, k, q);
          M(Q, k, p) = cs * x - sn * y;
          M(Q, k, q) = sn * x + cs * y;
        }
      }
  }
  int ord[24];
  for (int i = 0; i < n; i++) ord[i] = i;
  std::sort(ord, ord + n, [&](int x, int y) { return M(C, x, x) < M(C, y, y); });
  for (int c = 0; c < n; c++) {
    const int oc = ord[c];
    lam[c] = M(C, oc, oc);
    for (int i = n - 1; i >= 0; i--) {
      double s = M(Q, i, oc);
      for (int k = i + 1; k < n; k++) s -= M(L, k, i) * M(a, k, c);
      M(a, i, c) = s / M(L, i, i);
    }
  }
}

// hsmg.f:749-820 extended 1-D stiffness/mass, generalev, row_zero; writes
// S (nl x nl) and S^T into s[0..2*nl*nl)
void setup_fast1d(double *s, double *lam, int nl, int lbc, int rbc, double ll, double lm,
                  double lr, const double *ah, const double *bh, int n) {
  double a[24 * 24], b[24 * 24];
  auto A = [&](int i, int j) -> double & { return a[i + nl * j]; };
  auto B = [&](int i, int j) -> double & { return b[i + nl * j]; };
  auto AH = [&](int i, int j) { return ah[i + (n + 1) * j]; };
  const int i0 = (lbc == 1) ? 1 : 0, i1 = (rbc == 1) ? n - 1 : n;
  for (int q = 0; q < nl * nl; q++) a[q] = b[q] = 0.0;
  double fac = 2.0 / lm;
  A(1, 1) = 1.0;
  A(n + 1, n + 1) = 1.0;
  for (int j = i0; j <= i1; j++)
    for (int i = i0; i <= i1; i++) A(i + 1, j + 1) = fac * AH(i, j);
  if (lbc == 0) {
    fac = 2.0 / ll;
    A(0, 0) = fac * AH(n - 1, n - 1);
    A(1, 0) = fac * AH(n, n - 1);
    A(0, 1) = fac * AH(n - 1, n);
    A(1, 1) = A(1, 1) + fac * AH(n, n);
  } else {
    A(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 2.0 / lr;
    A(n + 1, n + 1) = A(n + 1, n + 1) + fac * AH(0, 0);
    A(n + 2, n + 1) = fac * AH(1, 0);
    A(n + 1, n + 2) = fac * AH(0, 1);
    A(n + 2, n + 2) = fac * AH(1, 1);
  } else {
    A(n + 2, n + 2) = 1.0;
  }
  fac = 0.5 * lm;
  B(1, 1) = 1.0;
  B(n + 1, n + 1) = 1.0;
  for (int i = i0; i <= i1; i++) B(i + 1, i + 1) = fac * bh[i];
  if (lbc == 0) {
    fac = 0.5 * ll;
    B(0, 0) = fac * bh[n - 1];
    B(1, 1) = B(1, 1) + fac * bh[n];
  } else {
    B(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 0.5 * lr;
    B(n + 1, n + 1) = B(n + 1, n + 1) + fac * bh[0];
    B(n + 2, n + 2) = fac * bh[1];
  } else {
    B(n + 2, n + 2) = 1.0;
  }
  generalev(a, b, lam, nl);
  auto zero_row = [&](int r) {
    for (int j = 0; j < nl; j++) a[r + nl * j] = 0.0;
  };
  if (lbc > 0) zero_row(0);
  if (lbc == 1) zero_row(1);
  if (rbc > 0) zero_row(nl - 1);
  if (rbc == 1) zero_row(nl - 2);
  for (int q = 0; q < nl * nl; q++) s[q] = a[q];
  for (int i = 0; i < nl; i++)
    for (int j = 0; j < nl; j++) s[nl * nl + j + nl * i] = a[i + nl * j];
}

}  // namespace mg1d

// GLL basis from capi.cpp
namespace spec {
void zwgll(double *z, double *w, int np);
}
namespace numbering {
int64_t setvert3d(int64_t *glo, int nx, int64_t nel, const int64_t *vtx);
}

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
namespace {

__global__ void k_fill_d(double *a, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// zero the nodes of faces with BC code 1 (connect1.f facev) at level nh
__global__ void k_face_zero(double *a, const int *fbc, int nh, int64_t nel) {
  const int64_t n3 = (int64_t)nh * nh * nh;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = l / n3;
    const int q = (int)(l - e * n3);
    const int i = q % nh, j = (q / nh) % nh, k = q / (nh * nh);
    const int *f = fbc + 6 * e;
    if ((f[0] == 1 && i == 0) || (f[1] == 1 && i == nh - 1) || (f[2] == 1 && j == 0) ||
        (f[3] == 1 && j == nh - 1) || (f[4] == 1 && k == 0) || (f[5] == 1 && k == nh - 1))
      a[l] = 0.0;
  }
}

// 1 on element-border nodes, 0 inside (hsmg.f:987-1067 restriction weights)
__global__ void k_border_ones(double *a, int nh, int64_t nel) {
  const int64_t n3 = (int64_t)nh * nh * nh;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(l % n3);
    const int i = q % nh, j = (q / nh) % nh, k = q / (nh * nh);
    a[l] = (i == 0 || i == nh - 1 || j == 0 || j == nh - 1 || k == 0 || k == nh - 1) ? 1.0 : 0.0;
  }
}

__global__ void k_rstr_wt(double *w, int nh, int64_t nel) {
  const int64_t n3 = (int64_t)nh * nh * nh;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(l % n3);
    const int i = q % nh, j = (q / nh) % nh, k = q / (nh * nh);
    const bool b = (i == 0 || i == nh - 1 || j == 0 || j == nh - 1 || k == 0 || k == nh - 1);
    w[l] = b ? 1.0 / w[l] : 1.0;
  }
}

// Schwarz weights: the loops of hsmg_schwarz_wt3d (hsmg.f:1285-1319) folded
// into one node array; a node visited twice (n = 3) is weighted twice.
__global__ void k_schwarz_wt(double *wt, const double *cnt, int n, int64_t nel) {
  const int64_t n3 = (int64_t)n * n * n;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(l % n3);
    const int i = q % n, j = (q / n) % n, k = q / (n * n);
    auto hits = [n](int a) { return (a == 0) + (a == 1) + (a == n - 2) + (a == n - 1); };
    const bool inx = (i >= 2 && i <= n - 3), iny = (j >= 2 && j <= n - 3);
    const int h = hits(i) + (inx ? hits(j) : 0) + ((inx && iny) ? hits(k) : 0);
    double w = 1.0;
    const double r = 1.0 / cnt[l];
    for (int t = 0; t < h; ++t) w *= r;
    wt[l] = w;
  }
}

// extended-array index
__device__ __forceinline__ int64_t xidx(int i, int j, int k, int m, int64_t e) {
  return i + m * (j + m * (k + (int64_t)m * e));
}

// W1 = ext(mask*r) with ghost faces = layer 2 values (extrude(0,0,2,1) of
// toext3d, hsmg.f:455-465); also r *= mask in place (h1mg_mask).
__global__ void k_sch_in(double *__restrict__ W1, double *__restrict__ r,
                         const double *__restrict__ mask, int n, int64_t nel) {
  const int m = n + 2;
  const int64_t m3 = (int64_t)m * m * m;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m3 * nel;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / m3;
    const int q = (int)(t - e * m3);
    const int i = q % m, j = (q / m) % m, k = q / (m * m);
    const int gi = (i == 0 || i == m - 1), gj = (j == 0 || j == m - 1), gk = (k == 0 || k == m - 1);
    double v = 0.0;
    if (gi + gj + gk == 0) {
      const int64_t l = (i - 1) + n * ((j - 1) + n * ((k - 1) + (int64_t)n * e));
      v = r[l] * mask[l];
      r[l] = v;
    } else if (gi + gj + gk == 1) {
      // ghost face node: value of the layer two planes inward (ext index 2 or m-3)
      int ii = i, jj = j, kk = k;
      if (gi) ii = (i == 0) ? 2 : m - 3;
      if (gj) jj = (j == 0) ? 2 : m - 3;
      if (gk) kk = (k == 0) ? 2 : m - 3;
      const int64_t l = (ii - 1) + n * ((jj - 1) + n * ((kk - 1) + (int64_t)n * e));
      v = r[l] * mask[l];
    }
    W1[t] = v;
  }
}

// FDM local solve on NL^3 blocks (hsmg.f:897-929 hsmg_do_fast):
//   W2 = (St (x) Ss (x) Sr) D (St^T (x) Ss^T (x) Sr^T) W1c
// with W1c = W1 whose ghost faces are reduced by layer 2 (the
// extrude(W1,0,1,W1,2,-1) of hsmg.f:467).  D(i,j,k) = 1/(lr_i + ls_j + lt_k)
// where that sum exceeds eps, else 0 (hsmg.f:753-770), is formed from the
// per-element eigenvalues instead of being stored.  The ghost faces of W2 are
// also written to G (the copy hsmg.f:476 keeps before the exchange).
// One element per block, NL*NL threads (i,j) with the k line in registers.
template <int NL>
__global__ void __launch_bounds__(NL *NL)
    k_fdm(double *__restrict__ W2, double *__restrict__ G, const double *__restrict__ W1,
          const double *__restrict__ S, const double *__restrict__ lam, int64_t nel) {
  constexpr int N2 = NL * NL, N3 = N2 * NL, NLAM = 3 * NL + 1;
  __shared__ double su[N3];
  __shared__ double sS[3][N2];
  __shared__ double sl[NLAM];
  const int tid = threadIdx.x, i = tid % NL, j = tid / NL;
  const bool gi = (i == 0 || i == NL - 1), gj = (j == 0 || j == NL - 1);
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    __syncthreads();
    const double *Se = S + (int64_t)e * N2 * 3;
    for (int t = tid; t < N2 * 3; t += blockDim.x) (&sS[0][0])[t] = Se[t];
    for (int t = tid; t < NLAM; t += blockDim.x) sl[t] = lam[(int64_t)e * NLAM + t];
    double v[NL], w[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const bool gk = (k == 0 || k == NL - 1);
      double x = W1[xidx(i, j, k, NL, e)];
      if ((int)gi + (int)gj + (int)gk == 1) {
        int ii = i, jj = j, kk = k;
        if (gi) ii = (i == 0) ? 2 : NL - 3;
        if (gj) jj = (j == 0) ? 2 : NL - 3;
        if (gk) kk = (k == 0) ? 2 : NL - 3;
        x = x - W1[xidx(ii, jj, kk, NL, e)];
      }
      su[i + NL * (j + NL * k)] = x;
    }
    __syncthreads();
    // r: w(i,j,k) = sum_m Sr(m,i) u(m,j,k)
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[0][mm + NL * i] * su[mm + NL * (j + NL * k)];
      w[k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; ++k) su[i + NL * (j + NL * k)] = w[k];
    __syncthreads();
    // s: w(i,j,k) = sum_m Ss(m,j) u(i,m,k)
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[1][mm + NL * j] * su[i + NL * (mm + NL * k)];
      w[k] = acc;
    }
    // t (registers) and the diagonal
    const double eps = sl[3 * NL], lij = sl[i] + sl[NL + j];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[2][mm + NL * k] * w[mm];
      const double diag = lij + sl[2 * NL + k];
      v[k] = (diag > eps) ? acc * (1.0 / diag) : 0.0;
    }
    // back: t (registers), s, r
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[2][k + NL * mm] * v[mm];
      w[k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; ++k) su[i + NL * (j + NL * k)] = w[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[1][j + NL * mm] * su[i + NL * (mm + NL * k)];
      v[k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; ++k) su[i + NL * (j + NL * k)] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int mm = 0; mm < NL; ++mm) acc += sS[0][i + NL * mm] * su[mm + NL * (j + NL * k)];
      const int64_t x = xidx(i, j, k, NL, e);
      W2[x] = acc;
      const bool gk = (k == 0 || k == NL - 1);
      if ((int)gi + (int)gj + (int)gk == 1) G[x] = acc;
    }
  }
}

// After the ext gather-scatter of W2: ghost -= G (pre-exchange ghost),
// layer 2 += ghost (extrude(0,1,0,-1) then extrude(2,1,0,1), hsmg.f:479-481),
// e = interior (toreg3d).  One thread per interior node.
__global__ void k_sch_out(double *__restrict__ e_out, const double *__restrict__ W2,
                          const double *__restrict__ G, int n, int64_t nel) {
  const int m = n + 2;
  const int64_t n3 = (int64_t)n * n * n;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = l / n3;
    const int q = (int)(l - e * n3);
    const int i = q % n + 1, j = (q / n) % n + 1, k = q / (n * n) + 1;  // ext coords 1..n
    double v = W2[xidx(i, j, k, m, e)];
    // the three extrude blocks run in order r, s, t; each adds the ghost of
    // its direction to layers 2 and m-3 over the face range 1..m-2
    auto ghost = [&](int gi, int gj, int gk) {
      const int64_t x = xidx(gi, gj, gk, m, e);
      return W2[x] - G[x];
    };
    if (i == 2) v += ghost(0, j, k);
    if (i == m - 3) v += ghost(m - 1, j, k);
    if (j == 2) v += ghost(i, 0, k);
    if (j == m - 3) v += ghost(i, m - 1, k);
    if (k == 2) v += ghost(i, j, 0);
    if (k == m - 3) v += ghost(i, j, m - 1);
    e_out[l] = v;
  }
}

// e = e * mask * wt * sigma
__global__ void k_sch_fin(double *__restrict__ e, const double *__restrict__ mask,
                          const double *__restrict__ wt, double sigma, int64_t n) {
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    double v = e[l] * mask[l];
    v = v * wt[l];
    e[l] = sigma * v;
  }
}

// per-element tensor apply with one matrix in all directions:
// v(NV^3) = (A (x) A (x) A) u(NU^3), A (NV x NU) column-major; optional
// pre-weight of u and accumulate into v.
template <int NV, int NU, bool ACC, bool PREW>
__global__ void k_tnsr(double *__restrict__ v, const double *__restrict__ u,
                       const double *__restrict__ wu, const double *__restrict__ A, int64_t nel) {
  constexpr int NV3 = NV * NV * NV, NU3 = NU * NU * NU;
  __shared__ double sA[NV * NU];
  __shared__ double su[NU3];
  __shared__ double s1[NV * NU * NU];
  __shared__ double s2[NV * NV * NU];
  for (int t = threadIdx.x; t < NV * NU; t += blockDim.x) sA[t] = A[t];
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    __syncthreads();
    for (int t = threadIdx.x; t < NU3; t += blockDim.x) {
      double x = u[(int64_t)e * NU3 + t];
      if constexpr (PREW) x = x * wu[(int64_t)e * NU3 + t];
      su[t] = x;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NV * NU * NU; t += blockDim.x) {
      const int i = t % NV, jk = t / NV;
      double s = 0.0;
      for (int m = 0; m < NU; ++m) s += sA[i + NV * m] * su[m + NU * jk];
      s1[t] = s;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NV * NV * NU; t += blockDim.x) {
      const int i = t % NV, j = (t / NV) % NV, k = t / (NV * NV);
      double s = 0.0;
      for (int m = 0; m < NU; ++m) s += s1[i + NV * (m + NU * k)] * sA[j + NV * m];
      s2[t] = s;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NV3; t += blockDim.x) {
      const int ij = t % (NV * NV), k = t / (NV * NV);
      double s = 0.0;
      for (int m = 0; m < NU; ++m) s += s2[ij + NV * NV * m] * sA[k + NV * m];
      if constexpr (ACC) v[(int64_t)e * NV3 + t] += s;
      else v[(int64_t)e * NV3 + t] = s;
    }
  }
}

// coarse: b[slot] = sum of the level-1 local values in that slot (CSR by slot)
__global__ void k_crs_gather(double *__restrict__ b, const double *__restrict__ r,
                             const int32_t *__restrict__ off, const int32_t *__restrict__ idx,
                             int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int q = off[s]; q < off[s + 1]; ++q) acc += r[idx[q]];
    b[s] = acc;
  }
}

__global__ void k_crs_scatter(double *__restrict__ e, const double *__restrict__ x,
                              const int32_t *__restrict__ slot, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int s = slot[q];
    e[q] = s >= 0 ? x[s] : 0.0;
  }
}

// w = the same nxyz-block in every element
__global__ void k_bcast_elem(double *__restrict__ w, const double *__restrict__ b, int64_t nxyz,
                             int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    w[q] = b[q % nxyz];
}

// Galerkin coarse element matrix column j (navier8.f:1648-1687):
// Ae(i,j,e) = b_i . (A_e b_j) for the 8 trilinear basis functions b_i.
__global__ void __launch_bounds__(256)
    k_galerkin_dot(double *__restrict__ Ae, const double *__restrict__ bas,
                   const double *__restrict__ w, int64_t nxyz, int64_t nel, int j) {
  __shared__ double sb[8][4];
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    double acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.0;
    for (int64_t q = threadIdx.x; q < nxyz; q += blockDim.x) {
      const double we = w[e * nxyz + q];
#pragma unroll
      for (int i = 0; i < 8; ++i) acc[i] += bas[i * nxyz + q] * we;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double v = acc[i];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if ((threadIdx.x & 63) == 0) sb[i][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
      const int i = threadIdx.x;
      Ae[64 * e + i + 8 * j] = ((sb[i][0] + sb[i][1]) + sb[i][2]) + sb[i][3];
    }
    __syncthreads();
  }
}

struct CrsCg {
  double rz, rz_old, pAp, alpha, beta, rr0, rr, tol2;
  int iter, done, maxit;
};

// x = 0, r = b, z = dinv*r, p = z; partial sums of r.z and r.r
__global__ void __launch_bounds__(256)
    k_ccg_init(double *x, double *r, double *p, const double *b, const double *dinv, int n,
               double *part) {
  __shared__ double s1[4], s2[4];
  double a = 0.0, c = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ri = b[i];
    x[i] = 0.0;
    r[i] = ri;
    const double zi = dinv[i] * ri;
    p[i] = zi;
    a += ri * zi;
    c += ri * ri;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    c += __shfl_down(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s1[threadIdx.x >> 6] = a;
    s2[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ((s1[0] + s1[1]) + s1[2]) + s1[3];
    part[2 * blockIdx.x + 1] = ((s2[0] + s2[1]) + s2[2]) + s2[3];
  }
}

__global__ void __launch_bounds__(256)
    k_ccg_start(CrsCg *st, const double *part, int nb, double rtol) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, c = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    a += part[2 * q];
    c += part[2 * q + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int q = 0; q < (int)blockDim.x; ++q) {
      x += s1[q];
      y += s2[q];
    }
    st->rz = x;
    st->rr0 = y;
    st->rr = y;
    st->tol2 = rtol * rtol * y;
    st->iter = 0;
    st->done = (y == 0.0);
  }
}

// w = A p (CSR), partial p.w
__global__ void __launch_bounds__(256)
    k_ccg_spmv(double *__restrict__ w, const double *__restrict__ p,
               const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
               const double *__restrict__ va, int n, const CrsCg *st, double *part) {
  __shared__ double s1[4];
  if (st->done) return;
  double a = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = rp[i]; q < rp[i + 1]; ++q) s += va[q] * p[ci[q]];
    w[i] = s;
    a += p[i] * s;
  }
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) s1[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((s1[0] + s1[1]) + s1[2]) + s1[3];
}

__global__ void __launch_bounds__(256)
    k_ccg_alpha(CrsCg *st, const double *part, int nb) {
  __shared__ double s1[256];
  if (st->done) return;
  double a = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) a += part[q];
  s1[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0;
    for (int q = 0; q < (int)blockDim.x; ++q) x += s1[q];
    st->pAp = x;
    st->alpha = st->rz / x;
  }
}

// x += alpha p ; r -= alpha w ; z = dinv r ; partial r.z, r.r
__global__ void __launch_bounds__(256)
    k_ccg_update(double *x, double *r, const double *p, const double *w, const double *dinv,
                 double *z, int n, const CrsCg *st, double *part) {
  __shared__ double s1[4], s2[4];
  if (st->done) return;
  const double al = st->alpha;
  double a = 0.0, c = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += al * p[i];
    const double ri = r[i] - al * w[i];
    r[i] = ri;
    const double zi = dinv[i] * ri;
    z[i] = zi;
    a += ri * zi;
    c += ri * ri;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    c += __shfl_down(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s1[threadIdx.x >> 6] = a;
    s2[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ((s1[0] + s1[1]) + s1[2]) + s1[3];
    part[2 * blockIdx.x + 1] = ((s2[0] + s2[1]) + s2[2]) + s2[3];
  }
}

__global__ void __launch_bounds__(256)
    k_ccg_beta(CrsCg *st, const double *part, int nb) {
  __shared__ double s1[256], s2[256];
  if (st->done) return;
  double a = 0.0, c = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    a += part[2 * q];
    c += part[2 * q + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int q = 0; q < (int)blockDim.x; ++q) {
      x += s1[q];
      y += s2[q];
    }
    st->rz_old = st->rz;
    st->rz = x;
    st->rr = y;
    st->beta = x / st->rz_old;
    st->iter += 1;
    if (y <= st->tol2 || st->iter >= st->maxit) st->done = 1;
  }
}

__global__ void k_ccg_p(double *p, const double *z, int n, const CrsCg *st) {
  if (st->done) return;
  const double b = st->beta;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = z[i] + b * p[i];
}

int grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

}  // namespace

// ---------------------------------------------------------------------------
// hierarchy
// ---------------------------------------------------------------------------
// Scalars of one hmh_flex_cg solve, resident on the device.
struct FcgState {
  double rho0, rho1, rho2, beta, den, alpha, rnorm, div0, tolpss;
  int iter, done;
};

struct MgLevel {
  int nx = 0, nh = 0;
  std::vector<double> ah, bh, zh;
  double *J = nullptr;  // level l -> l+1 interpolation (nh[l+1] x nh[l])
  double *Jt = nullptr; // its transpose (nh[l] x nh[l+1])
  GsMap gs, gsx;
  bool own_gs = false;
  double *mask = nullptr, *rwt = nullptr, *swt = nullptr;
  double *fs = nullptr, *fd = nullptr;
  double *e = nullptr, *r = nullptr;   // work vectors at this level
  double *W1 = nullptr, *W2 = nullptr, *G = nullptr;
  int64_t n = 0, nx3 = 0;
};

struct Mg {
  int lmax = 0;
  int64_t nel = 0;
  MgLevel lev[4];
  int *fbc = nullptr;  // device
  // coarse
  int ncrs = 0;
  int32_t *slot = nullptr;             // 8*nel
  int32_t *goff = nullptr, *gidx = nullptr;  // slot -> local entries
  int32_t *rp = nullptr, *ci = nullptr;
  double *va = nullptr, *dinv = nullptr;
  double *cx = nullptr, *cr = nullptr, *cp = nullptr, *cz = nullptr, *cw = nullptr, *cb = nullptr;
  double *cpart = nullptr;
  CrsCg *cst = nullptr;
  double crs_rtol = 1e-13;
  int crs_maxit = 4000;
  int crs_last_iters = 0;
  // flexible-CG vectors (5 fine fields) and scalars
  double *fr = nullptr, *fr1 = nullptr, *fp = nullptr, *fz = nullptr, *fw = nullptr;
  double *fpart = nullptr;
  FcgState *fst = nullptr;
  int64_t crs_iters_total = 0, crs_solves = 0;
};

static void mg_levels(int lx1, int *lmax, int *nx) {
  static const int mgn2[10] = {1, 2, 2, 2, 2, 3, 3, 5, 5, 5};
  const int lx2 = lx1;
  *lmax = (lx1 == 4) ? 2 : 3;
  nx[1] = 1;
  int mglx2 = 2 * (lx2 / 4) + 1;
  if (lx1 == 5) mglx2 = 3;
  if (lx1 <= 10) mglx2 = mgn2[std::min(lx1, 10) - 1];
  if (lx1 == 8) mglx2 = 4;
  if (lx1 == 8) mglx2 = 3;
  mglx2 = std::min(3, mglx2);
  nx[2] = mglx2;
  nx[3] = mglx2 + 1;
  nx[*lmax] = lx1 - 1;
}

// semhat (fast3d.f:1215-1292): ah = D^T B D with Fornberg derivatives
static void semhat(int n, std::vector<double> &ah, std::vector<double> &bh,
                   std::vector<double> &zh) {
  const int np = n + 1;
  ah.assign(np * np, 0.0);
  bh.assign(np, 0.0);
  zh.assign(np, 0.0);
  spec::zwgll(zh.data(), bh.data(), np);
  std::vector<double> d(np * np), w(2 * np);
  for (int i = 0; i <= n; i++) {
    mg1d::fd_weights_full(zh[i], zh.data(), n, 1, w.data());
    for (int j = 0; j <= n; j++) d[i + np * j] = w[j + np];
  }
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++) {
      double s = 0.0;
      for (int k = 0; k <= n; k++) s = s + d[k + np * i] * bh[k] * d[k + np * j];
      ah[i + np * j] = s;
    }
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src) {
  NK_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(src.size(), 1)));
  if (!src.empty())
    NK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return NKH_OK;
}

static int dalloc(double **p, int64_t n) {
  if (hipMalloc(p, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) {
    set_error("multigrid allocation of %lld doubles failed", (long long)n);
    return NKH_ERR_NOMEM;
  }
  return NKH_OK;
}

// the overlap sum of h1mg_schwarz_part1 from W2 (local solutions, with G =
// its ghost faces before the exchange) to level values e
static int overlap_sum(Mg &mg, int l, hipStream_t s) {
  MgLevel &L = mg.lev[l];
  NK_TRY(gs_op(L.gsx, L.W2, 0, s));
  hipLaunchKernelGGL(k_sch_out, dim3(grid_of(L.n)), dim3(256), 0, s, L.e, L.W2, L.G, L.nh,
                     mg.nel);
  NK_HIP(hipGetLastError());
  NK_TRY(gs_op(L.gs, L.e, 0, s));
  return NKH_OK;
}

template <int NL>
static void launch_fdm(MgLevel &L, int64_t nel, hipStream_t s) {
  const int grid = (int)std::min<int64_t>(nel, 1 << 20);
  hipLaunchKernelGGL(k_fdm<NL>, dim3(grid), dim3(NL * NL), 0, s, L.W2, L.G, L.W1, L.fs, L.fd,
                     nel);
}

static int run_fdm(MgLevel &L, int64_t nel, hipStream_t s) {
  switch (L.nh + 2) {
    case 5: launch_fdm<5>(L, nel, s); break;
    case 6: launch_fdm<6>(L, nel, s); break;
    case 7: launch_fdm<7>(L, nel, s); break;
    case 8: launch_fdm<8>(L, nel, s); break;
    case 9: launch_fdm<9>(L, nel, s); break;
    case 10: launch_fdm<10>(L, nel, s); break;
    case 11: launch_fdm<11>(L, nel, s); break;
    case 12: launch_fdm<12>(L, nel, s); break;
    case 13: launch_fdm<13>(L, nel, s); break;
    case 14: launch_fdm<14>(L, nel, s); break;
    case 15: launch_fdm<15>(L, nel, s); break;
    case 16: launch_fdm<16>(L, nel, s); break;
    case 17: launch_fdm<17>(L, nel, s); break;
    case 18: launch_fdm<18>(L, nel, s); break;
    default: set_error("FDM block size %d not instantiated", L.nh + 2); return NKH_ERR_UNSUPPORTED;
  }
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// tensor apply dispatch for (nv, nu) pairs used by the level transfers
template <bool ACC, bool PREW>
static int run_tnsr(double *v, int nv, const double *u, int nu, const double *wu,
                    const double *A, int64_t nel, hipStream_t s) {
  const int grid = (int)std::min<int64_t>(nel, 1 << 20);
#define NK_T(V, U)                                                                            \
  if (nv == V && nu == U) {                                                                   \
    hipLaunchKernelGGL((k_tnsr<V, U, ACC, PREW>), dim3(grid), dim3(V * V * V > 256 ? 256 : 64), \
                       0, s, v, u, wu, A, nel);                                               \
    NK_HIP(hipGetLastError());                                                                \
    return NKH_OK;                                                                            \
  }
  NK_T(2, 3) NK_T(3, 2) NK_T(2, 4) NK_T(4, 2) NK_T(3, 4) NK_T(4, 3) NK_T(4, 5) NK_T(5, 4)
  NK_T(4, 6) NK_T(6, 4) NK_T(4, 7) NK_T(7, 4) NK_T(4, 8) NK_T(8, 4) NK_T(4, 9) NK_T(9, 4)
  NK_T(4, 10) NK_T(10, 4) NK_T(4, 11) NK_T(11, 4) NK_T(4, 12) NK_T(12, 4) NK_T(4, 13)
  NK_T(13, 4) NK_T(4, 14) NK_T(14, 4) NK_T(4, 15) NK_T(15, 4) NK_T(4, 16) NK_T(16, 4)
  NK_T(3, 5) NK_T(5, 3)
#undef NK_T
  set_error("level transfer %d <-> %d not instantiated", nv, nu);
  return NKH_ERR_UNSUPPORTED;
}


// Schwarz at level l on r (masked in place), result in L.e (hsmg.f:425-494)
static int schwarz(Mg &mg, int l, double *r, double sigma, hipStream_t s) {
  MgLevel &L = mg.lev[l];
  hipLaunchKernelGGL(k_sch_in, dim3(grid_of(L.nx3)), dim3(256), 0, s, L.W1, r, L.mask, L.nh,
                     mg.nel);
  NK_HIP(hipGetLastError());
  NK_TRY(gs_op(L.gsx, L.W1, 0, s));
  NK_TRY(run_fdm(L, mg.nel, s));
  NK_TRY(overlap_sum(mg, l, s));
  hipLaunchKernelGGL(k_sch_fin, dim3(grid_of(L.n)), dim3(256), 0, s, L.e, L.mask, L.swt, sigma,
                     L.n);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

static int coarse_solve(Mg &mg, double *e1, const double *r1, hipStream_t s) {
  const int n = mg.ncrs;
  const int nb = std::min(grid_of(n), 1024);
  hipLaunchKernelGGL(k_crs_gather, dim3(grid_of(n)), dim3(256), 0, s, mg.cb, r1, mg.goff,
                     mg.gidx, n);
  hipLaunchKernelGGL(k_ccg_init, dim3(nb), dim3(256), 0, s, mg.cx, mg.cr, mg.cp, mg.cb, mg.dinv,
                     n, mg.cpart);
  hipLaunchKernelGGL(k_ccg_start, dim3(1), dim3(256), 0, s, mg.cst, mg.cpart, nb, mg.crs_rtol);
  NK_HIP(hipGetLastError());
  CrsCg hs{};
  int done_iters = 0;
  // the first batch covers the previous solve's count: the kernels return at
  // once after convergence, so over-enqueueing costs only empty launches
  int chunk = std::max(16, mg.crs_last_iters + 4);
  while (true) {
    for (int q = 0; q < chunk; ++q) {
      hipLaunchKernelGGL(k_ccg_spmv, dim3(nb), dim3(256), 0, s, mg.cw, mg.cp, mg.rp, mg.ci,
                         mg.va, n, mg.cst, mg.cpart);
      hipLaunchKernelGGL(k_ccg_alpha, dim3(1), dim3(256), 0, s, mg.cst, mg.cpart, nb);
      hipLaunchKernelGGL(k_ccg_update, dim3(nb), dim3(256), 0, s, mg.cx, mg.cr, mg.cp, mg.cw,
                         mg.dinv, mg.cz, n, mg.cst, mg.cpart);
      hipLaunchKernelGGL(k_ccg_beta, dim3(1), dim3(256), 0, s, mg.cst, mg.cpart, nb);
      hipLaunchKernelGGL(k_ccg_p, dim3(nb), dim3(256), 0, s, mg.cp, mg.cz, n, mg.cst);
    }
    NK_HIP(hipGetLastError());
    done_iters += chunk;
    NK_HIP(hipMemcpyAsync(&hs, mg.cst, sizeof(CrsCg), hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
    NK_CHECK(done_iters <= mg.crs_maxit + chunk, NKH_ERR_STATE, "coarse CG did not stop");
    chunk = 16;
  }
  mg.crs_last_iters = hs.iter;
  mg.crs_iters_total += hs.iter;
  mg.crs_solves += 1;
  hipLaunchKernelGGL(k_crs_scatter, dim3(grid_of(8 * mg.nel)), dim3(256), 0, s, e1, mg.cx,
                     mg.slot, 8 * mg.nel);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// hsmg.f:1855-1949 H1MG_SOLVE: z = M rhs (rhs masked in place).
int mg_solve(Mesh &mesh, double *z, double *rhs, hipStream_t s) {
  NK_CHECK(mesh.mg, NKH_ERR_STATE, "mesh has no multigrid hierarchy (nkh_mg_setup)");
  Mg &mg = *mesh.mg;
  Mesh *m = &mesh;
  const int Lm = mg.lmax;
  MgLevel &T = mg.lev[Lm];
  // fine Schwarz: result to z
  NK_TRY(schwarz(mg, Lm, rhs, 1.0, s));
  NK_HIP(hipMemcpyAsync(z, T.e, sizeof(double) * T.n, hipMemcpyDeviceToDevice, s));
  // downward leg: r_l = J^T (rwt_{l+1} r_{l+1}), dssum, Schwarz
  const double *rsrc = rhs;
  for (int l = Lm - 1; l >= 2; --l) {
    MgLevel &C = mg.lev[l], &F = mg.lev[l + 1];
    NK_TRY((run_tnsr<false, true>(C.r, C.nh, rsrc, F.nh, F.rwt, C.Jt, mg.nel, s)));
    NK_TRY(gs_op(C.gs, C.r, 0, s));
    NK_TRY(schwarz(mg, l, C.r, 1.0, s));
    rsrc = C.r;
  }
  // to level 1 without dssum, mask, coarse solve, mask
  MgLevel &L1 = mg.lev[1], &L2 = mg.lev[2];
  NK_TRY((run_tnsr<false, true>(L1.r, L1.nh, rsrc, L2.nh, L2.rwt, L1.Jt, mg.nel, s)));
  NK_TRY(mul_inplace(L1.r, L1.mask, L1.n, s));
  NK_TRY(coarse_solve(mg, L1.e, L1.r, s));
  NK_TRY(mul_inplace(L1.e, L1.mask, L1.n, s));
  // upward: e_l += J e_{l-1}; z += J e_{lmax-1}
  for (int l = 2; l <= Lm; ++l) {
    MgLevel &C = mg.lev[l - 1], &F = mg.lev[l];
    double *tgt = (l == Lm) ? z : F.e;
    NK_TRY((run_tnsr<true, false>(tgt, F.nh, C.e, C.nh, nullptr, C.J, mg.nel, s)));
  }
  // dsavg (ic.f:1871-1897): dssum then times vmult = 1/multiplicity
  NK_TRY(gs_op(m->gs, z, 4, s));
  return NKH_OK;
}

static void free_level(MgLevel &L) {
  for (double *p : {L.J, L.Jt, L.mask, L.rwt, L.swt, L.fs, L.fd, L.e, L.r, L.W1, L.W2, L.G})
    if (p) (void)hipFree(p);
  if (L.own_gs) gs_free(L.gs);
  gs_free(L.gsx);
}

void mg_free(Mesh &mesh) {
  Mg *mg = mesh.mg;
  if (!mg) return;
  for (int l = 1; l <= mg->lmax; ++l) free_level(mg->lev[l]);
  for (void *p : {(void *)mg->fbc, (void *)mg->slot, (void *)mg->goff, (void *)mg->gidx,
                  (void *)mg->rp, (void *)mg->ci, (void *)mg->va, (void *)mg->dinv, (void *)mg->cx,
                  (void *)mg->cr, (void *)mg->cp, (void *)mg->cz, (void *)mg->cw, (void *)mg->cb,
                  (void *)mg->cpart, (void *)mg->cst, (void *)mg->fr, (void *)mg->fr1,
                  (void *)mg->fp, (void *)mg->fz, (void *)mg->fw, (void *)mg->fpart,
                  (void *)mg->fst})
    if (p) (void)hipFree(p);
  delete mg;
  mesh.mg = nullptr;
}

// hsmg.f:2234-2270 H1MG_SETUP (+ fast3d.f swap_lengths, navier8.f
// set_up_h1_crs).  fbc: per element 6 BC codes in symmetric face order;
// x/y/z: fine GLL coordinates (host).
int mg_setup(Mesh &mesh, const int *fbc_h, const int64_t *vertex, const double *x,
             const double *y, const double *z) {
  Mesh *m = &mesh;
  NK_CHECK(m->has_geom && m->has_gs, NKH_ERR_STATE, "mesh not ready for multigrid");
  NK_CHECK(m->np >= 4, NKH_ERR_UNSUPPORTED, "multigrid needs lx1 >= 4 (got %d)", m->np);
  for (int64_t q = 0; q < 6 * m->nel; ++q)
    NK_CHECK(fbc_h[q] >= 0 && fbc_h[q] <= 2, NKH_ERR_ARG, "face BC code %d (element %lld)",
             fbc_h[q], (long long)(q / 6));
  hipStream_t s = cur_stream();
  mg_free(mesh);
  Mg *pmg = new Mg();
  mesh.mg = pmg;
  Mg &mg = *pmg;
  const int64_t nel = m->nel;
  mg.nel = nel;
  int nx[4] = {0, 0, 0, 0};
  mg_levels(m->np, &mg.lmax, nx);
  NK_TRY(upload(&mg.fbc, std::vector<int>(fbc_h, fbc_h + 6 * nel)));
  for (int l = 1; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    L.nx = nx[l];
    L.nh = nx[l] + 1;
    L.n = (int64_t)L.nh * L.nh * L.nh * nel;
    L.nx3 = (int64_t)(L.nh + 2) * (L.nh + 2) * (L.nh + 2) * nel;
    semhat(L.nx, L.ah, L.bh, L.zh);
  }
  for (int l = 1; l < mg.lmax; ++l) {  // hsmg.f:90-113 interpolation l -> l+1
    MgLevel &C = mg.lev[l], &F = mg.lev[l + 1];
    const int nf = F.nh, nc = C.nh;
    std::vector<double> J(nf * nc), Jt(nf * nc), w(2 * nc);
    for (int i = 0; i < nf; i++) {
      mg1d::fd_weights_full(F.zh[i], C.zh.data(), nc - 1, 1, w.data());
      for (int j = 0; j < nc; j++) {
        J[i + nf * j] = w[j];
        Jt[j + nc * i] = w[j];
      }
    }
    NK_TRY(upload(&C.J, J));
    NK_TRY(upload(&C.Jt, Jt));
  }
  // numbering and gs maps per level (hsmg.f:2361-2394)
  for (int l = 1; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    if (l == mg.lmax) {
      L.gs = m->gs;  // the fine map (setvert3d at lx1)
      L.own_gs = false;
    } else {
      std::vector<int64_t> glo((size_t)L.n);
      numbering::setvert3d(glo.data(), L.nh, nel, vertex);
      NK_TRY(gs_build(L.gs, glo.data(), L.n, L.nh, nel));
      L.own_gs = true;
    }
    if (l >= 2) {
      std::vector<int64_t> glox((size_t)L.nx3);
      numbering::setvert3d(glox.data(), L.nh + 2, nel, vertex);
      NK_TRY(gs_build(L.gsx, glox.data(), L.nx3, L.nh + 2, nel));
    }
  }
  // masks, restriction weights, work arrays
  for (int l = 1; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    NK_TRY(dalloc(&L.mask, L.n));
    NK_TRY(dalloc(&L.rwt, L.n));
    NK_TRY(dalloc(&L.e, L.n));
    NK_TRY(dalloc(&L.r, L.n));
    hipLaunchKernelGGL(k_fill_d, dim3(grid_of(L.n)), dim3(256), 0, s, L.mask, 1.0, L.n);
    hipLaunchKernelGGL(k_face_zero, dim3(grid_of(L.n)), dim3(256), 0, s, L.mask, mg.fbc, L.nh,
                       nel);
    NK_TRY(gs_op(L.gs, L.mask, 1, s));
    hipLaunchKernelGGL(k_border_ones, dim3(grid_of(L.n)), dim3(256), 0, s, L.rwt, L.nh, nel);
    NK_TRY(gs_op(L.gs, L.rwt, 0, s));
    hipLaunchKernelGGL(k_rstr_wt, dim3(grid_of(L.n)), dim3(256), 0, s, L.rwt, L.nh, nel);
    NK_HIP(hipGetLastError());
    if (l >= 2) {
      NK_TRY(dalloc(&L.W1, L.nx3));
      NK_TRY(dalloc(&L.W2, L.nx3));
      NK_TRY(dalloc(&L.G, L.nx3));
      NK_TRY(dalloc(&L.swt, L.n));
    }
  }
  // element lengths (fast3d.f:1542-1617 swap_lengths with plane_space
  // :306-423; the weight of face node j is the 1-based wxm1(j) as read there)
  const int n = m->np, n2 = n - 1, nxi = n - 2;
  std::vector<double> lm(3 * nel), ll(3 * nel), lr(3 * nel);
  {
    auto X = [&](const double *a, int i, int j, int k, int64_t e) {
      return a[i + n * (j + n * (k + (int64_t)n * e))];
    };
    const double *w = m->w;
    for (int64_t e = 0; e < nel; ++e) {
      for (int dir = 0; dir < 3; ++dir) {
        double sacc = 0.0, ws = 0.0;
        for (int b = 1; b <= nxi; ++b)
          for (int a = 1; a <= nxi; ++a) {
            int i0, j0, k0, i1, j1, k1;
            double wt;
            if (dir == 0) {
              i0 = 0; i1 = n2; j0 = j1 = a; k0 = k1 = b; wt = w[a - 1] * w[b - 1];
            } else if (dir == 1) {
              j0 = 0; j1 = n2; i0 = i1 = a; k0 = k1 = b; wt = w[a - 1] * w[b - 1];
            } else {
              k0 = 0; k1 = n2; i0 = i1 = a; j0 = j1 = b; wt = w[a - 1] * w[b - 1];
            }
            const double dx = X(x, i1, j1, k1, e) - X(x, i0, j0, k0, e);
            const double dy = X(y, i1, j1, k1, e) - X(y, i0, j0, k0, e);
            const double dz = X(z, i1, j1, k1, e) - X(z, i0, j0, k0, e);
            sacc = sacc + wt / (dx * dx + dy * dy + dz * dz);
            ws = ws + wt;
          }
        sacc = sacc / ws;
        lm[3 * e + dir] = 1.0 / std::sqrt(sacc);
      }
    }
    // neighbour lengths through dssum of lm written on face interiors
    std::vector<double> Lh((size_t)m->n, 0.0);
    for (int64_t e = 0; e < nel; ++e)
      for (int j = 1; j <= n - 2; j++)
        for (int k = 1; k <= n - 2; k++) {
          auto at = [&](int i, int jj, int kk) -> double & {
            return Lh[(size_t)(i + n * (jj + n * (kk + (int64_t)n * e)))];
          };
          at(0, k, j) = lm[3 * e];
          at(n - 1, k, j) = lm[3 * e];
          at(k, 0, j) = lm[3 * e + 1];
          at(k, n - 1, j) = lm[3 * e + 1];
          at(k, j, 0) = lm[3 * e + 2];
          at(k, j, n - 1) = lm[3 * e + 2];
        }
    double *d;
    NK_TRY(dalloc(&d, m->n));
    NK_HIP(hipMemcpyAsync(d, Lh.data(), sizeof(double) * m->n, hipMemcpyHostToDevice, s));
    NK_TRY(gs_op(m->gs, d, 0, s));
    NK_HIP(hipMemcpyAsync(Lh.data(), d, sizeof(double) * m->n, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    (void)hipFree(d);
    auto at = [&](int i, int j, int k, int64_t e) {
      return Lh[(size_t)(i + n * (j + n * (k + (int64_t)n * e)))];
    };
    for (int64_t e = 0; e < nel; ++e) {
      ll[3 * e + 0] = at(0, 1, 1, e) - lm[3 * e];
      lr[3 * e + 0] = at(n - 1, 1, 1, e) - lm[3 * e];
      ll[3 * e + 1] = at(1, 0, 1, e) - lm[3 * e + 1];
      lr[3 * e + 1] = at(1, n - 1, 1, e) - lm[3 * e + 1];
      ll[3 * e + 2] = at(1, 1, 0, e) - lm[3 * e + 2];
      lr[3 * e + 2] = at(1, 1, n - 1, e) - lm[3 * e + 2];
    }
  }
  // FDM per level (hsmg.f:633-665, 700-771), host threads over elements
  for (int l = 2; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    const int nl = L.nh + 2, ord = L.nx;
    const int nlam = 3 * nl + 1;
    std::vector<double> S((size_t)nl * nl * 3 * nel), Lam((size_t)nlam * nel);
    auto work = [&](int64_t e0, int64_t e1) {
      double lam[3][24], s2[2 * 24 * 24];
      for (int64_t e = e0; e < e1; ++e) {
        const int *bc = fbc_h + 6 * e;
        for (int dir = 0; dir < 3; ++dir) {
          mg1d::setup_fast1d(s2, lam[dir], nl, bc[2 * dir], bc[2 * dir + 1], ll[3 * e + dir],
                             lm[3 * e + dir], lr[3 * e + dir], L.ah.data(), L.bh.data(), ord);
          std::memcpy(S.data() + (size_t)nl * nl * (3 * e + dir), s2, sizeof(double) * nl * nl);
        }
        double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int dir = 0; dir < 3; ++dir)
          for (int i = 1; i <= nl - 2; ++i) mx[dir] = std::max(mx[dir], lam[dir][i]);
        double *le = Lam.data() + (size_t)nlam * e;
        for (int dir = 0; dir < 3; ++dir)
          for (int i = 0; i < nl; ++i) le[dir * nl + i] = lam[dir][i];
        le[3 * nl] = 1.e-5 * (mx[0] + mx[1] + mx[2]);
      }
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) {
      const int64_t a = nel * t / nth, b = nel * (t + 1) / nth;
      th.emplace_back(work, a, b);
    }
    for (auto &t : th) t.join();
    NK_TRY(upload(&L.fs, S));
    NK_TRY(upload(&L.fd, Lam));
  }
  // Schwarz weights: the overlap sum of ones (hsmg.f:3048-3104)
  for (int l = 2; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    hipLaunchKernelGGL(k_fill_d, dim3(grid_of(L.nx3)), dim3(256), 0, s, L.W2, 1.0, L.nx3);
    hipLaunchKernelGGL(k_fill_d, dim3(grid_of(L.nx3)), dim3(256), 0, s, L.G, 1.0, L.nx3);
    NK_TRY(overlap_sum(mg, l, s));
    hipLaunchKernelGGL(k_schwarz_wt, dim3(grid_of(L.n)), dim3(256), 0, s, L.swt, L.e, L.nh, nel);
    NK_HIP(hipGetLastError());
  }
  // coarse space: vertex ids, Dirichlet vertices dropped via the level-1 mask
  {
    MgLevel &L1 = mg.lev[1];
    std::vector<double> cm((size_t)L1.n);
    NK_HIP(hipMemcpyAsync(cm.data(), L1.mask, sizeof(double) * L1.n, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    int64_t vmax = 0;
    for (int64_t q = 0; q < 8 * nel; ++q) vmax = std::max(vmax, vertex[q]);
    std::vector<int32_t> slot_of((size_t)vmax + 1, -1), slot((size_t)(8 * nel));
    int ncrs = 0;
    for (int64_t q = 0; q < 8 * nel; ++q) {
      if (cm[(size_t)q] < 0.1) {
        slot[(size_t)q] = -1;
        continue;
      }
      int32_t &sv = slot_of[(size_t)vertex[q]];
      if (sv < 0) sv = ncrs++;
      slot[(size_t)q] = sv;
    }
    mg.ncrs = ncrs;
    std::vector<int32_t> goff(ncrs + 1, 0), gidx;
    for (int64_t q = 0; q < 8 * nel; ++q)
      if (slot[(size_t)q] >= 0) goff[slot[(size_t)q] + 1]++;
    for (int i = 0; i < ncrs; ++i) goff[i + 1] += goff[i];
    gidx.assign(goff[ncrs], 0);
    {
      std::vector<int32_t> pos(goff.begin(), goff.end() - 1);
      for (int64_t q = 0; q < 8 * nel; ++q)
        if (slot[(size_t)q] >= 0) gidx[pos[slot[(size_t)q]]++] = (int32_t)q;
    }
    NK_TRY(upload(&mg.slot, slot));
    NK_TRY(upload(&mg.goff, goff));
    NK_TRY(upload(&mg.gidx, gidx));
    // Galerkin element matrices b_i^T A_e b_j with the fine axhelm (h1=1,h2=0)
    const int64_t nxyz = (int64_t)n * n * n;
    std::vector<double> bas((size_t)(8 * nxyz));
    for (int j = 1; j <= 8; ++j) {
      double zr[kMaxNP], zs[kMaxNP], zt[kMaxNP];
      for (int i = 0; i < n; ++i) zr[i] = zs[i] = zt[i] = 0.5 * (1 - m->z[i]);
      if (j % 2 == 0)
        for (int i = 0; i < n; ++i) zr[i] = 0.5 * (1 + m->z[i]);
      if (j == 3 || j == 4 || j == 7 || j == 8)
        for (int i = 0; i < n; ++i) zs[i] = 0.5 * (1 + m->z[i]);
      if (j > 4)
        for (int i = 0; i < n; ++i) zt[i] = 0.5 * (1 + m->z[i]);
      for (int r = 0; r < n; ++r)
        for (int q = 0; q < n; ++q)
          for (int p = 0; p < n; ++p)
            bas[(size_t)((j - 1) * nxyz + p + n * (q + n * r))] = zr[p] * zs[q] * zt[r];
    }
    double *w1, *w2, *bd, *aed;
    NK_TRY(dalloc(&w1, m->n));
    NK_TRY(dalloc(&w2, m->n));
    NK_TRY(dalloc(&bd, 8 * nxyz));
    NK_TRY(dalloc(&aed, 64 * nel));
    NK_HIP(hipMemcpyAsync(bd, bas.data(), sizeof(double) * 8 * nxyz, hipMemcpyHostToDevice, s));
    Coef c1, c2;
    c1.s = 1.0;
    c2.s = 0.0;
    for (int j = 0; j < 8; ++j) {
      hipLaunchKernelGGL(k_bcast_elem, dim3(grid_of(m->n)), dim3(256), 0, s, w1, bd + j * nxyz,
                         nxyz, m->n);
      NK_HIP(hipGetLastError());
      NK_TRY(launch_axhelm(*m, w2, w1, c1, c2, s));
      hipLaunchKernelGGL(k_galerkin_dot, dim3((unsigned)std::min<int64_t>(nel, 1 << 20)),
                         dim3(256), 0, s, aed, bd, w2, nxyz, nel, j);
      NK_HIP(hipGetLastError());
    }
    std::vector<double> Ae((size_t)(64 * nel));
    NK_HIP(hipMemcpyAsync(Ae.data(), aed, sizeof(double) * 64 * nel, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    (void)hipFree(w1);
    (void)hipFree(w2);
    (void)hipFree(bd);
    (void)hipFree(aed);
    // assemble CSR (rows sorted, duplicates summed in element order)
    std::vector<std::vector<std::pair<int32_t, double>>> rows((size_t)ncrs);
    for (int64_t e = 0; e < nel; ++e)
      for (int j = 0; j < 8; ++j)
        for (int i = 0; i < 8; ++i) {
          const int32_t si = slot[(size_t)(8 * e + i)], sj = slot[(size_t)(8 * e + j)];
          if (si >= 0 && sj >= 0) rows[(size_t)si].emplace_back(sj, Ae[(size_t)(64 * e + i + 8 * j)]);
        }
    std::vector<int32_t> rp(ncrs + 1, 0), ci;
    std::vector<double> va, dg(ncrs, 0.0);
    for (int i = 0; i < ncrs; ++i) {
      auto &rw = rows[(size_t)i];
      std::stable_sort(rw.begin(), rw.end(),
                       [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
                         return a.first < b.first;
                       });
      for (size_t q = 0; q < rw.size();) {
        size_t t = q;
        double acc = 0.0;
        while (t < rw.size() && rw[t].first == rw[q].first) acc += rw[t++].second;
        ci.push_back(rw[q].first);
        va.push_back(acc);
        if (rw[q].first == i) dg[i] = acc;
        q = t;
      }
      rp[i + 1] = (int32_t)ci.size();
      rw.clear();
      rw.shrink_to_fit();
    }
    for (auto &v : dg) v = 1.0 / v;
    NK_TRY(upload(&mg.rp, rp));
    NK_TRY(upload(&mg.ci, ci));
    NK_TRY(upload(&mg.va, va));
    NK_TRY(upload(&mg.dinv, dg));
    for (double **p : {&mg.cx, &mg.cr, &mg.cp, &mg.cz, &mg.cw, &mg.cb}) NK_TRY(dalloc(p, ncrs));
    NK_TRY(dalloc(&mg.cpart, 2 * 1024));
    NK_HIP(hipMalloc(&mg.cst, sizeof(CrsCg)));
    CrsCg init{};
    init.maxit = mg.crs_maxit;
    NK_HIP(hipMemcpy(mg.cst, &init, sizeof(CrsCg), hipMemcpyHostToDevice));
  }
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

int mg_info(const Mesh &mesh, int *lmax, int *nh, int *ncrs, double *crs_avg_iters) {
  const Mg *mg = mesh.mg;
  NK_CHECK(mg, NKH_ERR_STATE, "mesh has no multigrid hierarchy");
  if (lmax) *lmax = mg->lmax;
  if (nh)
    for (int l = 1; l <= mg->lmax; ++l) nh[l - 1] = mg->lev[l].nh;
  if (ncrs) *ncrs = mg->ncrs;
  if (crs_avg_iters)
    *crs_avg_iters = mg->crs_solves ? (double)mg->crs_iters_total / mg->crs_solves : 0.0;
  return NKH_OK;
}

// ---------------------------------------------------------------------------
// hmh_flex_cg (hmholtz.f:2164-2281)
// ---------------------------------------------------------------------------
namespace {

constexpr int kFcgBlocks = 2048;  // fixed grid: deterministic partial sums

__device__ __forceinline__ void block_part(double a, double *part) {
  __shared__ double sa[4];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
}

__device__ __forceinline__ void block_part2(double a, double b, double *pa, double *pb) {
  __shared__ double sa[4], sb[4];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
    pb[blockIdx.x] = ((sb[0] + sb[1]) + sb[2]) + sb[3];
  }
}

// one block of 256 threads: fixed-order sum of kFcgBlocks partials
__device__ double final_sum(const double *part) {
  __shared__ double sf[4];
  double v = 0.0;
  for (int q = threadIdx.x; q < kFcgBlocks; q += blockDim.x) v += part[q];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sf[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = ((sf[0] + sf[1]) + sf[2]) + sf[3];
  __syncthreads();
  return t;
}

#define NK_GRID_LOOP(i, n)                                                     \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

// r = res; res = 0; r1 = 0; p = 0; partial glsc3(r, wt, r)
__global__ void __launch_bounds__(256)
    k_fcg_init(double *__restrict__ r, double *__restrict__ res, double *__restrict__ r1,
               double *__restrict__ p, const double *__restrict__ wt, int64_t n, double *part) {
  double a = 0.0;
  NK_GRID_LOOP(i, n) {
    const double ri = res[i];
    r[i] = ri;
    res[i] = 0.0;
    r1[i] = 0.0;
    p[i] = 0.0;
    a += (ri * wt[i]) * ri;
  }
  block_part(a, part);
}

__global__ void __launch_bounds__(256)
    k_fcg_start(FcgState *st, const double *part, double vol, double param21, double tolps) {
  const double t = final_sum(part);
  if (threadIdx.x == 0) {
    const double div0 = sqrt(t / vol);
    st->div0 = div0;
    st->tolpss = (param21 < 0) ? fabs(param21) * div0 : tolps;
    st->rho1 = 1.0;
    st->iter = 0;
    st->done = 0;
    st->rnorm = div0;
  }
}

// r1 <- r1 - r; partial glsc3(z,wt,r) and glsc3(z,wt,r1); r1 <- r
__global__ void __launch_bounds__(256)
    k_fcg_dots(const double *__restrict__ z, const double *__restrict__ r,
               double *__restrict__ r1, const double *__restrict__ wt, int64_t n, double *pa,
               double *pb) {
  double a = 0.0, b = 0.0;
  NK_GRID_LOOP(i, n) {
    const double ri = r[i], zw = z[i] * wt[i];
    const double d = r1[i] - ri;
    a += zw * ri;
    b += zw * d;
    r1[i] = ri;
  }
  block_part2(a, b, pa, pb);
}

__global__ void __launch_bounds__(256) k_fcg_beta(FcgState *st, const double *pa, const double *pb) {
  const double a = final_sum(pa);
  const double b = final_sum(pb);
  if (threadIdx.x == 0) {
    st->rho0 = st->rho1;
    st->rho1 = a;
    st->rho2 = -b;
    st->beta = st->rho2 / st->rho0;
  }
}

// add2s1: p = beta*p + z
__global__ void k_fcg_p(double *__restrict__ p, const double *__restrict__ z, const FcgState *st,
                        int64_t n) {
  const double beta = st->beta;
  NK_GRID_LOOP(i, n) p[i] = beta * p[i] + z[i];
}

// w = mask*w (after the dssum of ax); partial glsc3(w,wt,p)
__global__ void __launch_bounds__(256)
    k_fcg_den(double *__restrict__ w, const double *__restrict__ mask,
              const double *__restrict__ wt, const double *__restrict__ p, int64_t n,
              double *part) {
  double a = 0.0;
  NK_GRID_LOOP(i, n) {
    const double wi = w[i] * mask[i];
    w[i] = wi;
    a += (wi * wt[i]) * p[i];
  }
  block_part(a, part);
}

__global__ void __launch_bounds__(256) k_fcg_alpha(FcgState *st, const double *part) {
  const double t = final_sum(part);
  if (threadIdx.x == 0) {
    st->den = t;
    st->alpha = st->rho1 / t;
  }
}

// res += alpha p; r -= alpha w; partial sum r*r*wt
__global__ void __launch_bounds__(256)
    k_fcg_update(double *__restrict__ res, double *__restrict__ r, const double *__restrict__ p,
                 const double *__restrict__ w, const double *__restrict__ wt, const FcgState *st,
                 int64_t n, double *part) {
  const double alpha = st->alpha;
  double a = 0.0;
  NK_GRID_LOOP(i, n) {
    res[i] = res[i] + alpha * p[i];
    const double ri = r[i] - alpha * w[i];
    r[i] = ri;
    a = a + ri * ri * wt[i];
  }
  block_part(a, part);
}

__global__ void __launch_bounds__(256) k_fcg_rnorm(FcgState *st, const double *part, double vol) {
  const double t = final_sum(part);
  if (threadIdx.x == 0) {
    const double rn = sqrt(t / vol);
    st->rnorm = rn;
    st->iter += 1;
    if (rn < st->tolpss) st->done = 1;
  }
}

}  // namespace

int run_flex_cg(Mesh &m, double *res, Coef h1, Coef h2, const double *wt, const double *mask,
                int maxit, const nkh_params &prm, nkh_result *out, double *rnorm_history) {
  NK_CHECK(m.mg, NKH_ERR_STATE, "hmh_flex_cg needs the multigrid hierarchy (nkh_mg_setup)");
  Mg &mg = *m.mg;
  hipStream_t s = cur_stream();
  const int64_t n = m.n;
  if (!mg.fr) {
    for (double **p : {&mg.fr, &mg.fr1, &mg.fp, &mg.fz, &mg.fw}) NK_TRY(dalloc(p, n));
    NK_TRY(dalloc(&mg.fpart, 3 * kFcgBlocks));
    NK_HIP(hipMalloc(&mg.fst, sizeof(FcgState)));
  }
  double *r = mg.fr, *r1 = mg.fr1, *p = mg.fp, *z = mg.fz, *w = mg.fw;
  double *pa = mg.fpart, *pb = mg.fpart + kFcgBlocks;
  // hmholtz.f:2205-2208: tolps from chktcg1 (pmask, vmult), capped by param(21)
  double tolps = 0.0;
  NK_TRY(run_chktcg1(m, &tolps, res, h1, h2, mask, wt));
  if (prm.param21 > 0 && tolps > std::fabs(prm.param21)) tolps = std::fabs(prm.param21);
  if (prm.istep == 0) tolps = 1.e-4;
  const dim3 g(kFcgBlocks), b(256);
  hipLaunchKernelGGL(k_fcg_init, g, b, 0, s, r, res, r1, p, wt, n, pa);
  hipLaunchKernelGGL(k_fcg_start, dim3(1), b, 0, s, mg.fst, pa, m.vol, prm.param21, tolps);
  NK_HIP(hipGetLastError());
  Timers &tm = timers();
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  if (tm.on)
    for (auto &e : ev) NK_HIP(hipEventCreate(&e));
  FcgState hs{};
  int iter = 0;
  int rc = NKH_OK;
  for (int k = 1; k <= maxit; ++k) {
    if ((rc = mg_solve(m, z, r, s)) != NKH_OK) break;
    hipLaunchKernelGGL(k_fcg_dots, g, b, 0, s, z, r, r1, wt, n, pa, pb);
    hipLaunchKernelGGL(k_fcg_beta, dim3(1), b, 0, s, mg.fst, pa, pb);
    hipLaunchKernelGGL(k_fcg_p, dim3(grid_of(n)), b, 0, s, p, z, mg.fst, n);
    if (tm.on) (void)hipEventRecord(ev[0], s);
    if ((rc = launch_axhelm(m, w, p, h1, h2, s)) != NKH_OK) break;
    if (tm.on) (void)hipEventRecord(ev[1], s);
    if ((rc = gs_op(m.gs, w, 0, s)) != NKH_OK) break;
    if (tm.on) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(k_fcg_den, g, b, 0, s, w, mask, wt, p, n, pa);
    hipLaunchKernelGGL(k_fcg_alpha, dim3(1), b, 0, s, mg.fst, pa);
    hipLaunchKernelGGL(k_fcg_update, g, b, 0, s, res, r, p, w, wt, mg.fst, n, pa);
    hipLaunchKernelGGL(k_fcg_rnorm, dim3(1), b, 0, s, mg.fst, pa, m.vol);
    if (hipGetLastError() != hipSuccess) {
      set_error("hmh_flex_cg: kernel launch failed");
      rc = NKH_ERR_HIP;
      break;
    }
    if (hipMemcpyAsync(&hs, mg.fst, sizeof(FcgState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("hmh_flex_cg: state read-back failed");
      rc = NKH_ERR_HIP;
      break;
    }
    if (tm.on) {
      float a = 0.f, c = 0.f;
      (void)hipEventElapsedTime(&a, ev[0], ev[1]);
      (void)hipEventElapsedTime(&c, ev[1], ev[2]);
      tm.taxhm += a * 1e-3;
      tm.naxhm += 1;
      tm.tdsum += c * 1e-3;
      tm.ndsum += 1;
    }
    iter = hs.iter;
    if (rnorm_history) rnorm_history[k - 1] = hs.rnorm;
    if (hs.done) break;
  }
  if (tm.on)
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  NK_TRY(rc);
  if (iter == 0) {
    NK_HIP(hipMemcpyAsync(&hs, mg.fst, sizeof(FcgState), hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
  }
  if (out) {
    out->niter = iter;
    out->converged = hs.done;
    out->rbn0 = hs.div0;
    out->rbn2 = hs.rnorm;
    out->tol = hs.tolpss;
  }
  return NKH_OK;
}

}  // namespace nekamd
