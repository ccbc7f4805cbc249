// Set-up of the H1 multigrid hierarchy (hsmg.f:2234-2270 h1mg_setup and the
// routines it calls; fast3d.f:1542-1617 swap_lengths; navier8.f:83-233
// set_up_h1_crs), host side with device kernels for the per-node weights and
// the Galerkin coarse matrices.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hsmg.h"

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

// GLL basis and numbering from capi.cpp
namespace spec {
void zwgll(double *z, double *w, int np);
}
namespace numbering {
int64_t setvert3d(int64_t *glo, int nx, int64_t nel, const int64_t *vtx);
}

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

// mask * Schwarz weight.  The weight folds the loops of hsmg_schwarz_wt3d
// (hsmg.f:1285-1319) into one node array: each statement of those loops
// multiplies its node by 1/count, so a node the loops visit twice (n = 3) is
// weighted twice.  cnt = the overlap count (h1mg_setup_schwarz_wt_1,
// hsmg.f:3048-3104).
__global__ void k_mask_schwarz_wt(double *mswt, const double *mask, const double *cnt, int n,
                                  int64_t nel) {
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
    mswt[l] = mask[l] * w;
  }
}

// Byte codes of the fine level's mask * Schwarz weights: 0 where masked,
// else 1 + 8 (count - 1) + visits, the table entry formed as
// k_mask_schwarz_wt forms the weight; and of the restriction weights, the
// count c with 1/c the weight.  A node whose weight has no exact code sets
// bit 1 (mask * Schwarz) or bit 2 (restriction) of *bad.
__global__ void k_weight_codes(uint8_t *mc, uint8_t *rc, const double *mswt, const double *rwt,
                               const double *mask, const double *cnt, const double *lut, int n,
                               int64_t nel, int *bad) {
  const int64_t n3 = (int64_t)n * n * n;
  int b = 0;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n3 * nel;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(l % n3);
    const int i = q % n, j = (q / n) % n, k = q / (n * n);
    auto hits = [n](int a) { return (a == 0) + (a == 1) + (a == n - 2) + (a == n - 1); };
    const bool inx = (i >= 2 && i <= n - 3), iny = (j >= 2 && j <= n - 3);
    const int h = hits(i) + (inx ? hits(j) : 0) + ((inx && iny) ? hits(k) : 0);
    const double c = cnt[l];
    int code = 0;
    if (mask[l] != 0.0) {
      if (c >= 1.0 && c <= 31.0 && h <= 7) code = 1 + 8 * ((int)c - 1) + h;
      else b = 1;
    }
    mc[l] = (uint8_t)code;
    if (lut[code] != mswt[l]) b |= 1;
    // restriction weight: 1/count
    const double rw = rwt[l];
    const double rr = rw > 0.0 ? rint(1.0 / rw) : 0.0;
    const int rcode = (rr >= 1.0 && rr <= 255.0) ? (int)rr : 0;
    rc[l] = (uint8_t)rcode;
    if (rcode == 0 || 1.0 / (double)rcode != rw) b |= 2;
  }
  if (b) atomicOr(bad, b);
}

// 1 at ghost slots that have a neighbour (the overlap count of a field of
// ones), 0 elsewhere
__global__ void k_ghost_ones(double *w, const int32_t *gout, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = gout[i] >= 0 ? 1.0 : 0.0;
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

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  if (hipMalloc(dst, sizeof(T) * std::max<size_t>(src.size(), 1)) != hipSuccess) {
    set_error("multigrid allocation of %zu bytes failed", sizeof(T) * src.size());
    return NKH_ERR_NOMEM;
  }
  if (!src.empty())
    NK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return NKH_OK;
}

// hsmg.f:2272-2323 H1MG_SETUP_MG_NX (PnPn: lx2 = lx1)
void mg_levels(int lx1, int *lmax, int *nx) {
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
void semhat(int n, std::vector<double> &ah, std::vector<double> &bh, std::vector<double> &zh) {
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

// Ghost-face slot (f,a,b) of an element -> coordinates in its (n+2)^3
// extension (ghost layer at 0 or n+1).
inline void ghost_ext(int f, int a, int b, int n, int &i, int &j, int &k) {
  const int m1 = n + 1;
  switch (f) {
    case 0: i = 0; j = a + 1; k = b + 1; break;
    case 1: i = m1; j = a + 1; k = b + 1; break;
    case 2: i = a + 1; j = 0; k = b + 1; break;
    case 3: i = a + 1; j = m1; k = b + 1; break;
    case 4: i = a + 1; j = b + 1; k = 0; break;
    default: i = a + 1; j = b + 1; k = m1; break;
  }
}

// The node one layer inside face f of an element, at face position (a,b),
// in level-field coordinates (the layer the overlap exchange sends).
inline void inner_node(int f, int a, int b, int n, int &i, int &j, int &k) {
  const int in = (f & 1) ? n - 2 : 1;
  switch (f >> 1) {
    case 0: i = in; j = a; k = b; break;
    case 1: i = a; j = in; k = b; break;
    default: i = a; j = b; k = in; break;
  }
}

// Pair the ghost faces of neighbouring extended elements through the
// extended numbering (setvert3d with nh+2 points): the two ghost nodes that
// share an id face each other.
int build_ghost_maps(MgLevel &L, int64_t nel, const int64_t *vertex) {
  const int n = L.nh, m = n + 2;
  const int64_t n2 = (int64_t)n * n, nslot = 6 * n2 * nel;
  NK_CHECK(nslot < ((int64_t)1 << 31), NKH_ERR_ARG, "ghost slots exceed int32 indexing");
  const int64_t m3 = (int64_t)m * m * m;
  std::vector<int64_t> glox((size_t)(m3 * nel));
  const int64_t ngx = numbering::setvert3d(glox.data(), m, nel, vertex);
  std::vector<int32_t> partner((size_t)nslot, -1);
  {
    std::vector<int32_t> first((size_t)ngx + 1, -1);
    for (int64_t e = 0; e < nel; ++e)
      for (int f = 0; f < 6; ++f)
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a) {
            int i, j, k;
            ghost_ext(f, a, b, n, i, j, k);
            const int64_t id = glox[(size_t)(i + m * (j + m * (k + m * e)))];
            const int32_t q = (int32_t)((e * 6 + f) * n2 + a + n * b);
            NK_CHECK(id > 0 && id <= ngx, NKH_ERR_STATE, "extended numbering: ghost id %lld",
                     (long long)id);
            int32_t &p = first[(size_t)id];
            if (p < 0) {
              p = q;
            } else {
              NK_CHECK(partner[(size_t)p] < 0, NKH_ERR_ARG,
                       "a face is shared by more than two elements (element %lld)",
                       (long long)e);
              partner[(size_t)p] = q;
              partner[(size_t)q] = p;
            }
          }
  }
  std::vector<int32_t> gin((size_t)nslot, -1);
  for (int64_t q = 0; q < nslot; ++q) {
    const int32_t p = partner[(size_t)q];
    if (p < 0) continue;
    const int64_t e2 = p / (6 * n2);
    const int f2 = (int)((p / n2) % 6);
    const int r = (int)(p % n2), a2 = r % n, b2 = r / n;
    int i, j, k;
    inner_node(f2, a2, b2, n, i, j, k);
    gin[(size_t)q] = (int32_t)(i + n * (j + n * (k + n * e2)));
  }
  NK_TRY(upload(&L.gin, gin));
  NK_TRY(upload(&L.gout, partner));
  return NKH_OK;
}

}  // namespace

int mg_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

int mg_dalloc(double **p, int64_t n) {
  if (hipMalloc(p, sizeof(double) * std::max<int64_t>(n, 1)) != hipSuccess) {
    set_error("multigrid allocation of %lld doubles failed", (long long)n);
    return NKH_ERR_NOMEM;
  }
  return NKH_OK;
}

namespace {
__global__ void k_to_f32(float *o, const double *a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    o[i] = (float)a[i];
}
__global__ void k_fill_f(float *a, float v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}
int falloc(float **p, int64_t n) {
  if (hipMalloc(p, sizeof(float) * std::max<int64_t>(n, 1)) != hipSuccess) {
    set_error("multigrid allocation of %lld floats failed", (long long)n);
    return NKH_ERR_NOMEM;
  }
  return NKH_OK;
}
}  // namespace

// single-precision smoother data of level l: S, eigenvalues, local solution
// and ghost slots
static int prepare_level_fp32(Mg &mg, MgLevel &L, hipStream_t s) {
  if (L.fs32) return NKH_OK;  // this level is prepared (all four arrays or none)
  const int nl = L.nh + 2;
  const int64_t ns = (int64_t)nl * nl * 3 * mg.nel, nlam = (int64_t)(3 * nl + 1) * mg.nel;
  const int64_t ngh = (int64_t)6 * L.nh * L.nh * mg.nel;
  float *fs = nullptr, *fl = nullptr, *wi = nullptr, *wg = nullptr;
  int rc = falloc(&fs, ns);
  if (rc == NKH_OK) rc = falloc(&fl, nlam);
  if (rc == NKH_OK) rc = falloc(&wi, L.n);
  if (rc == NKH_OK) rc = falloc(&wg, ngh);
  if (rc != NKH_OK) {
    for (float *p : {fs, fl, wi, wg})
      if (p) (void)hipFree(p);
    return rc;
  }
  L.fs32 = fs;
  L.fl32 = fl;
  L.wint32 = wi;
  L.wgh32 = wg;
  hipLaunchKernelGGL(k_to_f32, dim3(mg_grid(ns)), dim3(256), 0, s, L.fs32, L.fs, ns);
  hipLaunchKernelGGL(k_to_f32, dim3(mg_grid(nlam)), dim3(256), 0, s, L.fl32, L.fl, nlam);
  // ghost slots without a neighbour are never written: 0
  hipLaunchKernelGGL(k_fill_f, dim3(mg_grid(ngh)), dim3(256), 0, s, L.wgh32, 0.0f, ngh);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// Single-precision smoother data: the fine level's (precision 1 and 2) and,
// for the mixed solver's inner V-cycles (inner), the intermediate levels'.
// Each level is prepared once.
int mg_prepare_fp32(Mg &mg, hipStream_t s, bool inner) {
  MgLevel &L = mg.lev[mg.lmax];
  if (inner)
    for (int l = 2; l < mg.lmax; ++l) NK_TRY(prepare_level_fp32(mg, mg.lev[l], s));
  NK_TRY(prepare_level_fp32(mg, L, s));
  if (L.mswt32 && L.rwt32) return NKH_OK;
  if (!L.mswt32) NK_TRY(falloc(&L.mswt32, L.n));
  if (!L.rwt32) NK_TRY(falloc(&L.rwt32, L.n));
  hipLaunchKernelGGL(k_to_f32, dim3(mg_grid(L.n)), dim3(256), 0, s, L.mswt32, L.mswt, L.n);
  hipLaunchKernelGGL(k_to_f32, dim3(mg_grid(L.n)), dim3(256), 0, s, L.rwt32, L.rwt, L.n);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

static void free_level(MgLevel &L) {
  for (double *p : {L.J, L.Jt, L.mask, L.rwt, L.mswt, L.fs, L.fl, L.wint, L.wgh, L.e, L.r})
    if (p) (void)hipFree(p);
  for (int32_t *p : {L.gin, L.gout})
    if (p) (void)hipFree(p);
  for (float *p : {L.fs32, L.fl32, L.wint32, L.wgh32, L.e32, L.mswt32, L.rwt32})
    if (p) (void)hipFree(p);
  for (uint8_t *p : {L.mswt8, L.rwt8})
    if (p) (void)hipFree(p);
  if (L.own_gs) gs_free(L.gs);
}

void mg_free(Mesh &mesh) {
  Mg *mg = mesh.mg;
  if (!mg) return;
  if (mg->sb) (void)hipStreamSynchronize(mg->sb);
  if (mg->sc) (void)hipStreamSynchronize(mg->sc);
  for (int l = 1; l <= mg->lmax; ++l) free_level(mg->lev[l]);
  mg->crs.release();
  for (void *p : {(void *)mg->fbc, (void *)mg->slot, (void *)mg->goff, (void *)mg->gidx,
                  (void *)mg->cb, (void *)mg->cx, (void *)mg->fr, (void *)mg->fp,
                  (void *)mg->fz, (void *)mg->fw, (void *)mg->fpart, (void *)mg->fst,
                  (void *)mg->fflag, (void *)mg->gV, (void *)mg->gZ, (void *)mg->gx,
                  (void *)mg->gr, (void *)mg->gw, (void *)mg->gpart, (void *)mg->gsm,
                  (void *)mg->xr, (void *)mg->xp, (void *)mg->xz, (void *)mg->xw,
                  (void *)mg->xx, (void *)mg->xwt, (void *)mg->xpart, mg->xst,
                  (void *)mg->opart, (void *)mg->wt8, (void *)mg->wlut})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : mg->tev)
    if (e) (void)hipEventDestroy(e);
  if (mg->ev_fork) (void)hipEventDestroy(mg->ev_fork);
  if (mg->ev_join) (void)hipEventDestroy(mg->ev_join);
  for (int l = 0; l < 4; ++l) {
    if (mg->ev_lr[l]) (void)hipEventDestroy(mg->ev_lr[l]);
    if (mg->ev_le[l]) (void)hipEventDestroy(mg->ev_le[l]);
  }
  if (mg->sb) (void)hipStreamDestroy(mg->sb);
  if (mg->sc) (void)hipStreamDestroy(mg->sc);
  delete mg;
  mesh.mg = nullptr;
}

// hsmg.f:2234-2270 H1MG_SETUP (with swap_lengths, fast3d.f:1542-1617, and
// the coarse problem of set_up_h1_crs, navier8.f:83-233).  fbc: per element
// 6 BC codes in symmetric face order; vertex: 8 ids per element; x/y/z: fine
// GLL coordinates (host).
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
  NK_HIP(hipStreamCreateWithFlags(&mg.sb, hipStreamNonBlocking));
  NK_HIP(hipStreamCreateWithFlags(&mg.sc, hipStreamNonBlocking));
  for (int l = 0; l < 4; ++l) {
    NK_HIP(hipEventCreateWithFlags(&mg.ev_lr[l], hipEventDisableTiming));
    NK_HIP(hipEventCreateWithFlags(&mg.ev_le[l], hipEventDisableTiming));
  }
  NK_HIP(hipEventCreateWithFlags(&mg.ev_fork, hipEventDisableTiming));
  NK_HIP(hipEventCreateWithFlags(&mg.ev_join, hipEventDisableTiming));
  for (hipEvent_t &e : mg.tev) NK_HIP(hipEventCreate(&e));
  int nx[4] = {0, 0, 0, 0};
  mg_levels(m->np, &mg.lmax, nx);
  NK_TRY(upload(&mg.fbc, std::vector<int>(fbc_h, fbc_h + 6 * nel)));
  mg.null_space = true;
  for (int64_t q = 0; q < 6 * nel; ++q)
    if (fbc_h[q] == 1) mg.null_space = false;
  for (int l = 1; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    L.nx = nx[l];
    L.nh = nx[l] + 1;
    L.n = (int64_t)L.nh * L.nh * L.nh * nel;
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
  // numbering and gs maps per level (hsmg.f:2361-2394); the extended
  // numbering becomes the ghost-face maps
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
    if (l >= 2) NK_TRY(build_ghost_maps(L, nel, vertex));
  }
  // masks, restriction weights, work arrays
  for (int l = 1; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    NK_TRY(mg_dalloc(&L.mask, L.n));
    NK_TRY(mg_dalloc(&L.rwt, L.n));
    NK_TRY(mg_dalloc(&L.e, L.n));
    NK_TRY(mg_dalloc(&L.r, L.n));
    hipLaunchKernelGGL(k_fill_d, dim3(mg_grid(L.n)), dim3(256), 0, s, L.mask, 1.0, L.n);
    hipLaunchKernelGGL(k_face_zero, dim3(mg_grid(L.n)), dim3(256), 0, s, L.mask, mg.fbc, L.nh,
                       nel);
    NK_TRY(gs_op(L.gs, L.mask, 1, s));
    hipLaunchKernelGGL(k_border_ones, dim3(mg_grid(L.n)), dim3(256), 0, s, L.rwt, L.nh, nel);
    NK_TRY(gs_op(L.gs, L.rwt, 0, s));
    hipLaunchKernelGGL(k_rstr_wt, dim3(mg_grid(L.n)), dim3(256), 0, s, L.rwt, L.nh, nel);
    NK_HIP(hipGetLastError());
    if (l >= 2) {
      const int64_t nn = (int64_t)L.nh * L.nh;
      NK_TRY(mg_dalloc(&L.wint, L.n));
      NK_TRY(mg_dalloc(&L.wgh, 6 * nn * nel));
      NK_TRY(mg_dalloc(&L.mswt, L.n));
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
            const double wt = w[a - 1] * w[b - 1];
            if (dir == 0) {
              i0 = 0; i1 = n2; j0 = j1 = a; k0 = k1 = b;
            } else if (dir == 1) {
              j0 = 0; j1 = n2; i0 = i1 = a; k0 = k1 = b;
            } else {
              k0 = 0; k1 = n2; i0 = i1 = a; j0 = j1 = b;
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
    NK_TRY(mg_dalloc(&d, m->n));
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
  // FDM per level (hsmg.f:633-665, 700-771), host threads over elements:
  // per element S_r, S_s, S_t and the eigenvalues with the eps of the guard
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
    NK_TRY(upload(&L.fl, Lam));
  }
  // Schwarz weights: the overlap count of ones (hsmg.f:3048-3104), then
  // mask * weight
  for (int l = 2; l <= mg.lmax; ++l) {
    MgLevel &L = mg.lev[l];
    const int64_t nn = (int64_t)L.nh * L.nh;
    hipLaunchKernelGGL(k_fill_d, dim3(mg_grid(L.n)), dim3(256), 0, s, L.wint, 1.0, L.n);
    hipLaunchKernelGGL(k_ghost_ones, dim3(mg_grid(6 * nn * nel)), dim3(256), 0, s, L.wgh,
                       L.gout, 6 * nn * nel);
    NK_TRY(mg_overlap_out(mg, l, L.e, s));
    // slots without a neighbour are never written by the solves: keep them 0
    hipLaunchKernelGGL(k_fill_d, dim3(mg_grid(6 * nn * nel)), dim3(256), 0, s, L.wgh, 0.0,
                       6 * nn * nel);
    NK_TRY(gs_op(L.gs, L.e, 0, s));
    hipLaunchKernelGGL(k_mask_schwarz_wt, dim3(mg_grid(L.n)), dim3(256), 0, s, L.mswt, L.mask,
                       L.e, L.nh, nel);
    NK_HIP(hipGetLastError());
    if (l == mg.lmax) {  // L.e holds the overlap counts here
      std::vector<double> lut(512, 0.0);
      for (int c = 1; c <= 31; ++c)
        for (int h = 0; h <= 7; ++h) {
          double w = 1.0;
          const double r = 1.0 / c;
          for (int t = 0; t < h; ++t) w *= r;
          lut[(size_t)(1 + 8 * (c - 1) + h)] = 1.0 * w;
        }
      // entries 256-511: the same weights rounded to FP32, as the
      // single-precision smoother's copies hold them
      for (int c = 0; c < 256; ++c) lut[(size_t)(256 + c)] = (double)(float)lut[(size_t)c];
      NK_TRY(upload(&mg.wlut, lut));
      NK_HIP(hipMalloc(&L.mswt8, (size_t)L.n));
      int *bad = nullptr, hbad = 1;
      NK_HIP(hipMalloc(&bad, sizeof(int)));
      NK_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
      NK_HIP(hipMalloc(&L.rwt8, (size_t)L.n));
      hipLaunchKernelGGL(k_weight_codes, dim3(mg_grid(L.n)), dim3(256), 0, s, L.mswt8, L.rwt8,
                         L.mswt, L.rwt, L.mask, L.e, mg.wlut, L.nh, nel, bad);
      NK_HIP(hipGetLastError());
      NK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
      NK_HIP(hipStreamSynchronize(s));
      (void)hipFree(bad);
      if (hbad & 2) {
        (void)hipFree(L.rwt8);
        L.rwt8 = nullptr;
      }
      if (hbad & 1) {
        (void)hipFree(L.mswt8);
        L.mswt8 = nullptr;
      }
    }
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
    NK_TRY(mg_dalloc(&w1, m->n));
    NK_TRY(mg_dalloc(&w2, m->n));
    NK_TRY(mg_dalloc(&bd, 8 * nxyz));
    NK_TRY(mg_dalloc(&aed, 64 * nel));
    NK_HIP(hipMemcpyAsync(bd, bas.data(), sizeof(double) * 8 * nxyz, hipMemcpyHostToDevice, s));
    Coef c1, c2;
    c1.s = 1.0;
    c2.s = 0.0;
    for (int j = 0; j < 8; ++j) {
      hipLaunchKernelGGL(k_bcast_elem, dim3(mg_grid(m->n)), dim3(256), 0, s, w1, bd + j * nxyz,
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
    // assemble (duplicates summed in element order, as the dense sum of the
    // exact solver) into CSR with sorted rows
    std::vector<std::vector<std::pair<int32_t, double>>> rows((size_t)ncrs);
    for (int64_t e = 0; e < nel; ++e)
      for (int j = 0; j < 8; ++j)
        for (int i = 0; i < 8; ++i) {
          const int32_t si = slot[(size_t)(8 * e + i)], sj = slot[(size_t)(8 * e + j)];
          if (si >= 0 && sj >= 0)
            rows[(size_t)si].emplace_back(sj, Ae[(size_t)(64 * e + i + 8 * j)]);
        }
    std::vector<int32_t> rp(ncrs + 1, 0), ci;
    std::vector<double> va;
    // null space: the last unknown held at zero (its row and column dropped,
    // unit diagonal), as the oracle's dense factor does
    const int32_t pin = (mg.null_space && ncrs > 0) ? ncrs - 1 : -1;
    for (int i = 0; i < ncrs; ++i) {
      auto &rw = rows[(size_t)i];
      if (pin >= 0) {
        if (i == pin) {
          rw.assign(1, std::make_pair(pin, 1.0));
        } else {
          rw.erase(std::remove_if(rw.begin(), rw.end(),
                                  [pin](const std::pair<int32_t, double> &a) {
                                    return a.first == pin;
                                  }),
                   rw.end());
        }
      }
      std::stable_sort(rw.begin(), rw.end(),
                       [](const std::pair<int32_t, double> &a,
                          const std::pair<int32_t, double> &b) { return a.first < b.first; });
      for (size_t q = 0; q < rw.size();) {
        size_t t = q;
        double acc = 0.0;
        while (t < rw.size() && rw[t].first == rw[q].first) acc += rw[t++].second;
        ci.push_back(rw[q].first);
        va.push_back(acc);
        q = t;
      }
      rp[i + 1] = (int32_t)ci.size();
      std::vector<std::pair<int32_t, double>>().swap(rw);
    }
    NK_TRY(mg.crs.setup(rp, ci, va, s));
    NK_TRY(mg_dalloc(&mg.cb, ncrs));
    NK_TRY(mg_dalloc(&mg.cx, ncrs));
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
    *crs_avg_iters = mg->crs.solves ? (double)mg->crs.iters_total / mg->crs.solves : 0.0;
  return NKH_OK;
}

}  // namespace nekamd
