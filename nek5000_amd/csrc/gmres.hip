// Right-preconditioned restarted GMRES with the h1mg V-cycle (gmres.f:304-545
// hmh_gmres), the reference's default pressure solver (param(42) = 0).
//
// Krylov vectors V (m+1) and preconditioned vectors Z (m) stay in HBM; the
// small Hessenberg problem (Givens rotations, back substitution) runs on the
// host from the Gram-Schmidt coefficients read back each step.  Dots are
// fixed-grid partial sums reduced in a fixed order.  ml = mu = 1 in the
// reference (uzawa_gmres_split, gmres.f:252-271), so there is no scaling.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hsmg.h"

namespace nekamd {

namespace {

constexpr int kGmBlocks = 2048;
constexpr int kDotChunk = 8;

#define NK_LOOP(i, n)                                                          \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ void store_part(double a, double *part) {
  __shared__ double sa[4];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
  __syncthreads();
}

// w = v * mgmask (h1mg_solve masks its right-hand side in place)
__global__ void k_gm_copy_mask(double *__restrict__ w, const double *__restrict__ v,
                               const double *__restrict__ mask, int64_t n) {
  NK_LOOP(i, n) w[i] = v[i] * mask[i];
}

// r = res - mask*w; partial sum r*r*wt   (r = res - A x with ax's mask)
__global__ void __launch_bounds__(256)
    k_gm_resid(double *__restrict__ r, const double *__restrict__ res,
               const double *__restrict__ w, const double *__restrict__ mask,
               const double *__restrict__ wt, int64_t n, int first, double *part) {
  double a = 0.0;
  NK_LOOP(i, n) {
    const double ri = first ? res[i] : res[i] - w[i] * mask[i];
    r[i] = ri;
    a += (ri * ri) * wt[i];
  }
  store_part(a, part);
}

// w *= mask (ax's pmask), then partial sums of (w*v_i)*wt for i in
// [i0, i0+cnt) (vlsc3, gmres.f:466)
__global__ void __launch_bounds__(256)
    k_gm_dots(double *__restrict__ w, const double *__restrict__ mask, int apply_mask,
              const double *__restrict__ V, int64_t n, int i0, int cnt,
              const double *__restrict__ wt, double *part) {
  double acc[kDotChunk];
#pragma unroll
  for (int c = 0; c < kDotChunk; ++c) acc[c] = 0.0;
  NK_LOOP(q, n) {
    double wq = w[q];
    if (apply_mask) {
      wq = wq * mask[q];
      w[q] = wq;
    }
    const double wv = wt[q];
#pragma unroll
    for (int c = 0; c < kDotChunk; ++c)
      if (c < cnt) acc[c] += (wq * V[(int64_t)(i0 + c) * n + q]) * wv;
  }
#pragma unroll
  for (int c = 0; c < kDotChunk; ++c)
    if (c < cnt) store_part(acc[c], part + (int64_t)(i0 + c) * kGmBlocks);
}

// out[i] = fixed-order sum of the kGmBlocks partials of dot i (block i)
__global__ void __launch_bounds__(256) k_gm_final(const double *part, double *out) {
  __shared__ double sa[4];
  const double *p = part + (int64_t)blockIdx.x * kGmBlocks;
  double a = 0.0;
  for (int q = threadIdx.x; q < kGmBlocks; q += blockDim.x) a += p[q];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
}

// w = w - h_i v_i for i = 0..j-1 in order (add2s2), then partial w*w*wt
__global__ void __launch_bounds__(256)
    k_gm_orth(double *__restrict__ w, const double *__restrict__ V,
              const double *__restrict__ h, int j, const double *__restrict__ wt, int64_t n,
              double *part) {
  double a = 0.0;
  NK_LOOP(q, n) {
    double wq = w[q];
    for (int i = 0; i < j; ++i) wq = wq + (-h[i]) * V[(int64_t)i * n + q];
    w[q] = wq;
    a += (wq * wq) * wt[q];
  }
  store_part(a, part);
}

// v = w * s (cmult2)
__global__ void k_gm_scale(double *__restrict__ v, const double *__restrict__ w, double s,
                           int64_t n) {
  NK_LOOP(i, n) v[i] = w[i] * s;
}

// x = x + c_i z_i for i = 0..j-1 in order
__global__ void k_gm_xupd(double *__restrict__ x, const double *__restrict__ Z,
                          const double *__restrict__ c, int j, int64_t n) {
  NK_LOOP(q, n) {
    double xq = x[q];
    for (int i = 0; i < j; ++i) xq = xq + c[i] * Z[(int64_t)i * n + q];
    x[q] = xq;
  }
}

__global__ void k_gm_zero(double *__restrict__ x, int64_t n) {
  NK_LOOP(i, n) x[i] = 0.0;
}

}  // namespace

int run_gmres(Mesh &m, double *res, Coef h1, Coef h2, const double *wt, const double *mask,
              int maxit, const nkh_params &prm, nkh_result *out, double *rnorm_history) {
  NK_CHECK(m.mg, NKH_ERR_STATE, "hmh_gmres needs the multigrid hierarchy (nkh_mg_setup)");
  NK_CHECK(prm.param40 == 0 || prm.param40 == 1, NKH_ERR_UNSUPPORTED,
           "hmh_gmres: coarse solver param40=%d (0 exact, 1 one AMG V-cycle)", prm.param40);
  Mg &mg = *m.mg;
  hipStream_t s = cur_stream();
  const int64_t n = m.n;
  const int M = kGmresM;
  if (!mg.gV) {
    NK_TRY(mg_dalloc(&mg.gV, (int64_t)(M + 1) * n));
    NK_TRY(mg_dalloc(&mg.gZ, (int64_t)M * n));
    for (double **p : {&mg.gx, &mg.gr, &mg.gw}) NK_TRY(mg_dalloc(p, n));
    NK_TRY(mg_dalloc(&mg.gpart, (int64_t)(M + 2) * kGmBlocks));
    NK_TRY(mg_dalloc(&mg.gsm, 2 * (M + 2)));
  }
  double *V = mg.gV, *Z = mg.gZ, *x = mg.gx, *r = mg.gr, *w = mg.gw, *part = mg.gpart;
  double *dh = mg.gsm, *dc = mg.gsm + (M + 2);
  const double *mgmask = mg.lev[mg.lmax].mask;
  const dim3 G(kGmBlocks), B(256);
  const double norm_fac = 1.0 / std::sqrt(m.vol);
  const bool nullsp = mg_null_space_on(mg, h2);
  // hmholtz.f-style tolerance (gmres.f:352-356)
  double tolps = 0.0;
  NK_TRY(run_chktcg1(m, &tolps, res, h1, h2, mask, wt));
  if (prm.param21 > 0 && tolps > std::fabs(prm.param21)) tolps = std::fabs(prm.param21);
  if (prm.istep == 0) tolps = 1.e-4;
  double tolpss = tolps;
  std::vector<double> H((size_t)M * M, 0.0), gam(M + 1, 0.0), c(M, 0.0), sn(M, 0.0), hv(M + 2);
  auto HH = [&](int i, int jj) -> double & { return H[(size_t)i + (size_t)M * jj]; };
  auto dot_final = [&](int count, double *host) -> int {
    hipLaunchKernelGGL(k_gm_final, dim3(count), B, 0, s, part, dh);
    NK_HIP(hipGetLastError());
    NK_HIP(hipMemcpyAsync(host, dh, sizeof(double) * count, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    return NKH_OK;
  };
  hipLaunchKernelGGL(k_gm_zero, dim3(mg_grid(n)), B, 0, s, x, n);
  double div0 = 0.0, rnorm = 0.0;
  int iter = 0, j = 0;
  bool iconv = false;
  while (!iconv) {
    // r = res (first) or res - A x
    if (iter > 0) {
      NK_TRY(launch_axhelm(m, w, x, h1, h2, s));
      NK_TRY(gs_op(m.gs, w, 0, s));
    }
    hipLaunchKernelGGL(k_gm_resid, G, B, 0, s, r, res, w, mask, wt, n, iter == 0 ? 1 : 0, part);
    double g2;
    NK_TRY(dot_final(1, &g2));
    gam[0] = std::sqrt(g2);
    if (iter == 0) {
      div0 = gam[0] * norm_fac;
      if (prm.param21 < 0) tolpss = std::fabs(prm.param21) * div0;
    }
    rnorm = 0.0;
    if (gam[0] == 0.0) break;
    hipLaunchKernelGGL(k_gm_scale, dim3(mg_grid(n)), B, 0, s, V, r, 1.0 / gam[0], n);
    for (j = 1; j <= M; ++j) {
      iter += 1;
      double *vj = V + (int64_t)(j - 1) * n, *zj = Z + (int64_t)(j - 1) * n;
      hipLaunchKernelGGL(k_gm_copy_mask, dim3(mg_grid(n)), B, 0, s, w, vj, mgmask, n);
      NK_TRY(mg_solve(m, zj, w, s, prm.param40, prm.precision));
      if (nullsp) NK_TRY(mg_ortho(mg, zj, n, s));  // gmres.f:427
      NK_TRY(launch_axhelm(m, w, zj, h1, h2, s));
      NK_TRY(gs_op(m.gs, w, 0, s));
      // classical Gram-Schmidt, one pass (gmres.f:464-474)
      for (int i0 = 0; i0 < j; i0 += kDotChunk) {
        const int cnt = std::min(kDotChunk, j - i0);
        hipLaunchKernelGGL(k_gm_dots, G, B, 0, s, w, mask, i0 == 0 ? 1 : 0, V, n, i0, cnt, wt,
                           part);
      }
      NK_HIP(hipGetLastError());
      NK_TRY(dot_final(j, hv.data()));
      for (int i = 1; i <= j; ++i) HH(i - 1, j - 1) = hv[(size_t)(i - 1)];
      // dh still holds h(1..j) on the device
      hipLaunchKernelGGL(k_gm_orth, G, B, 0, s, w, V, dh, j, wt, n, part);
      double a2;
      NK_TRY(dot_final(1, &a2));
      for (int i = 1; i <= j - 1; ++i) {  // previous Givens rotations
        const double temp = HH(i - 1, j - 1);
        HH(i - 1, j - 1) = c[i - 1] * temp + sn[i - 1] * HH(i, j - 1);
        HH(i, j - 1) = -sn[i - 1] * temp + c[i - 1] * HH(i, j - 1);
      }
      const double alpha = std::sqrt(a2);
      rnorm = 0.0;
      if (alpha == 0.0) {
        iconv = true;
        break;
      }
      const double l = std::sqrt(HH(j - 1, j - 1) * HH(j - 1, j - 1) + alpha * alpha);
      const double temp = 1.0 / l;
      c[j - 1] = HH(j - 1, j - 1) * temp;
      sn[j - 1] = alpha * temp;
      HH(j - 1, j - 1) = l;
      gam[j] = -sn[j - 1] * gam[j - 1];
      gam[j - 1] = c[j - 1] * gam[j - 1];
      rnorm = std::fabs(gam[j]) * norm_fac;
      if (rnorm_history) rnorm_history[iter - 1] = rnorm;
      if (iter + 1 > maxit || rnorm < tolpss) {
        iconv = true;
        break;
      }
      if (j == M) break;  // restart
      hipLaunchKernelGGL(k_gm_scale, dim3(mg_grid(n)), B, 0, s, V + (int64_t)j * n, w,
                         1.0 / alpha, n);
    }
    if (j > M) j = M;
    // back substitution c = H^-1 gamma; x += sum c_i z_i (gmres.f:518-529)
    for (int k = j; k >= 1; --k) {
      double temp = gam[k - 1];
      for (int i = j; i >= k + 1; --i) temp = temp - HH(k - 1, i - 1) * c[i - 1];
      c[k - 1] = temp / HH(k - 1, k - 1);
    }
    NK_HIP(hipMemcpyAsync(dc, c.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gm_xupd, dim3(mg_grid(n)), B, 0, s, x, Z, dc, j, n);
    NK_HIP(hipGetLastError());
  }
  // res = x (copy, gmres.f:535), orthogonalised against the null space (:534)
  NK_HIP(hipMemcpyAsync(res, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (nullsp) NK_TRY(mg_ortho(mg, res, n, s));
  NK_HIP(hipStreamSynchronize(s));
  if (out) {
    out->niter = iter;
    out->converged = rnorm < tolpss || gam[0] == 0.0;
    out->rbn0 = div0;
    out->rbn2 = rnorm;
    out->tol = tolpss;
  }
  return NKH_OK;
}

}  // namespace nekamd
