// Geometric factors and Jacobi diagonal kernels.
#include "element_kernels.h"

namespace nekamd {

// ---------------------------------------------------------------------------
// Geometric factors: coef.f:879-927 XYZRST (x_r = D x, x_s = x D^T per layer,
// x_t = x D^T over layers), coef.f:555-632 GLMAPM1 (Jacobian, cofactors via
// ascol5) and coef.f:633-785 GEODAT1 (G_ij = cof_i.cof_j / J * w3, bm1 = J*w3).
// ---------------------------------------------------------------------------
template <int NP>
__global__ void __launch_bounds__(NP * NP * epb<NP>())
    k_geom(DMat<NP> Dm, const double *__restrict__ Dg, const double *__restrict__ wg,
           const double *__restrict__ x, const double *__restrict__ y,
           const double *__restrict__ z, double *__restrict__ g, int64_t gstride,
           double *__restrict__ bm1, int64_t nel) {
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, EPB = epb<NP>();
  __shared__ double s_D[NP][NP + 1];
  __shared__ double s_x[EPB][3][NP3];
  const int tid = threadIdx.x;
  const int le = tid / NP2;
  const int ij = tid - le * NP2;
  const int i = ij % NP, j = ij / NP;
  for (int t = tid; t < NP2; t += blockDim.x) s_D[t % NP][t / NP] = Dg[t];
  for (int64_t eb = blockIdx.x; eb * EPB < nel; eb += gridDim.x) {
    const int64_t e = eb * EPB + le;
    const bool valid = (le < EPB) && (e < nel);
    const int64_t base = e * NP3 + ij;
    double cx[NP], cy[NP], cz[NP];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double a = 0.0, b = 0.0, c = 0.0;
      if (valid) {
        a = x[base + k * NP2];
        b = y[base + k * NP2];
        c = z[base + k * NP2];
      }
      cx[k] = a;
      cy[k] = b;
      cz[k] = c;
      s_x[le][0][k * NP2 + ij] = a;
      s_x[le][1][k * NP2 + ij] = b;
      s_x[le][2][k * NP2 + ij] = c;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double xr = 0, yr = 0, zr = 0, xs = 0, ys = 0, zs = 0, xt = 0, yt = 0, zt = 0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        const double dr = s_D[i][m], ds = s_D[j][m], dt = Dm.d[k + NP * m];
        xr += dr * s_x[le][0][k * NP2 + j * NP + m];
        yr += dr * s_x[le][1][k * NP2 + j * NP + m];
        zr += dr * s_x[le][2][k * NP2 + j * NP + m];
        xs += ds * s_x[le][0][k * NP2 + m * NP + i];
        ys += ds * s_x[le][1][k * NP2 + m * NP + i];
        zs += ds * s_x[le][2][k * NP2 + m * NP + i];
        xt += dt * cx[m];
        yt += dt * cy[m];
        zt += dt * cz[m];
      }
      if (!valid) continue;
      double jac = 0.0;
      jac = jac + xr * ys * zt;
      jac = jac + xt * yr * zs;
      jac = jac + xs * yt * zr;
      jac = jac - xr * yt * zs;
      jac = jac - xs * yr * zt;
      jac = jac - xt * ys * zr;
      const double rx = ys * zt - yt * zs, ry = xt * zs - xs * zt, rz = xs * yt - xt * ys;
      const double sx = yt * zr - yr * zt, sy = xr * zt - xt * zr, sz = xt * yr - xr * yt;
      const double tx = yr * zs - ys * zr, ty = xs * zr - xr * zs, tz = xr * ys - xs * yr;
      const double wj = 1.0 / jac;
      const double w3 = wg[i] * wg[j] * wg[k];
      const int64_t l = base + k * NP2;
      g[l] = (rx * rx + ry * ry + rz * rz) * wj * w3;
      g[gstride + l] = (sx * sx + sy * sy + sz * sz) * wj * w3;
      g[2 * gstride + l] = (tx * tx + ty * ty + tz * tz) * wj * w3;
      g[3 * gstride + l] = (rx * sx + ry * sy + rz * sz) * wj * w3;
      g[4 * gstride + l] = (rx * tx + ry * ty + rz * tz) * wj * w3;
      g[5 * gstride + l] = (sx * tx + sy * ty + sz * tz) * wj * w3;
      bm1[l] = jac * w3;
    }
  }
}

template <int NP>
static int geom_dispatch(Mesh &m, const double *x, const double *y, const double *z,
                         hipStream_t s) {
  DMat<NP> Dm;
  for (int t = 0; t < NP * NP; ++t) Dm.d[t] = m.D[t];
  constexpr int EPB = epb<NP>();
  int64_t ngroups = (m.nel + EPB - 1) / EPB;
  int grid = (int)(ngroups < ((int64_t)1 << 20) ? ngroups : ((int64_t)1 << 20));
  if (grid < 1) grid = 1;
  // weights: stash after D in the device basis buffer
  hipLaunchKernelGGL((k_geom<NP>), dim3(grid), dim3(NP * NP * EPB), 0, s, Dm, m.Dd,
                     m.Dd + kMaxNP * kMaxNP, x, y, z, m.g, m.n, m.bm1, m.nel);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int launch_geom(Mesh &m, const double *x, const double *y, const double *z, hipStream_t s) {
#define CALL(NP) geom_dispatch<NP>(m, x, y, z, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

// ---------------------------------------------------------------------------
// Jacobi diagonal, hmholtz.f:380-524 SETPREC (3-D, every element deformed):
//   d(ix,iy,iz) = sum_q g1(q,iy,iz) D(q,ix)^2 + g2(ix,q,iz) D(q,iy)^2
//               + g3(ix,iy,q) D(q,iz)^2
// plus the g4/g5/g6 cross terms the reference adds at the 8 element corners
// (loops 600/700/800 step by lx1-1), then d = d*h1 + h2*bm1.  The dssum and
// inversion that follow are done by the caller.
// ---------------------------------------------------------------------------
template <int NP>
__global__ void k_setprec(DMat<NP> Dm, const double *__restrict__ g, int64_t gstride,
                          const double *__restrict__ h1f, double h1s,
                          const double *__restrict__ h2f, double h2s,
                          const double *__restrict__ bm1, double *__restrict__ d, int64_t n) {
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, L = NP - 1;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = l / NP3;
    const int q = (int)(l - e * NP3);
    const int ix = q % NP, iy = (q / NP) % NP, iz = q / NP2;
    const double *g1 = g + e * NP3, *g2 = g1 + gstride, *g3 = g1 + 2 * gstride,
                 *g4 = g1 + 3 * gstride, *g5 = g1 + 4 * gstride, *g6 = g1 + 5 * gstride;
    // DXTM1(a,b) = D(b,a) = Dm.d[b + NP*a]
#define DT(a, b) Dm.d[(b) + NP * (a)]
    double v = 0.0;
    for (int iq = 0; iq < NP; ++iq) v += g1[iq + NP * iy + NP2 * iz] * (DT(ix, iq) * DT(ix, iq));
    for (int iq = 0; iq < NP; ++iq) v += g2[ix + NP * iq + NP2 * iz] * (DT(iy, iq) * DT(iy, iq));
    for (int iq = 0; iq < NP; ++iq) v += g3[ix + NP * iy + NP2 * iq] * (DT(iz, iq) * DT(iz, iq));
    const bool ex = (ix == 0 || ix == L), ey = (iy == 0 || iy == L), ez = (iz == 0 || iz == L);
    if (ex && ey && ez) {
      // loop 600 (x-ends), 700 (y-ends), 800 (z-ends), in that order
      v = v + g4[q] * DT(ix, ix) * DT(iy, iy) + g5[q] * DT(ix, ix) * DT(iz, iz);
      v = v + g4[q] * DT(iy, iy) * DT(ix, ix) + g6[q] * DT(iy, iy) * DT(iz, iz);
      v = v + g5[q] * DT(iz, iz) * DT(ix, ix) + g6[q] * DT(iz, iz) * DT(iy, iy);
    }
#undef DT
    const double h1 = h1f ? h1f[l] : h1s;
    const double h2 = h2f ? h2f[l] : h2s;
    v = v * h1;
    v = v + h2 * bm1[l];
    d[l] = v;
  }
}

int launch_setprec(const Mesh &m, double *d, Coef h1, Coef h2, hipStream_t s) {
  const int64_t n = m.n;
  int grid = (int)((n + 255) / 256);
  if (grid > 65536) grid = 65536;
  if (grid < 1) grid = 1;
#define CALL(NP)                                                                      \
  ([&]() -> int {                                                                     \
    DMat<NP> Dm;                                                                      \
    for (int t = 0; t < NP * NP; ++t) Dm.d[t] = m.D[t];                               \
    hipLaunchKernelGGL((k_setprec<NP>), dim3(grid), dim3(256), 0, s, Dm, m.g, m.n,    \
                       h1.f, h1.s, h2.f, h2.s, m.bm1, d, n);                          \
    NK_HIP(hipGetLastError());                                                        \
    return NKH_OK;                                                                    \
  })()
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

// Undeformed elements (genxyz.f:351-470 setdef; param(59) = 0): the
// reference's axhelm and setprec skip their cross terms g4..g6
// (hmholtz.f:188-218 fast path, `if (ifdfrm(e))`; :440-468), so they are
// held as zeros here, whatever sfastax (hmholtz.f:353-364) left in Nek's
// g4m1..g6m1.
__global__ void k_zero_cross(double *__restrict__ g, const uint8_t *__restrict__ undeformed,
                             int64_t npe, int64_t nel, int64_t n) {
  const int64_t total = 3 * n;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t l = q % n;
    if (undeformed[l / npe]) g[3 * n + q] = 0.0;
  }
}

int apply_undeformed(Mesh &m, hipStream_t s) {
  if (!m.undeformed || !m.g) return NKH_OK;
  const int64_t npe = (int64_t)m.np * m.np * m.np;
  int64_t grid = (3 * m.n + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_zero_cross, dim3((int)grid), dim3(256), 0, s, m.g, m.undeformed, npe,
                     m.nel, m.n);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

}  // namespace nekamd
