// Adjoint coefficient gradient of the Helmholtz solve.  For J(u) with
// A(h1,h2) u = f and the adjoint A lam = dJ/du (A is symmetric, so the
// forward solver serves for lam), the nodal gradients are
//   dJ/dh1_l = -(D lam)_l^T G_l (D u)_l,   dJ/dh2_l = -lam_l bm1_l u_l,
// with D the reference gradient (local_grad3, navier5.f:428-445) and G the
// six geometric factors paired as axhelm pairs them (hmholtz.f:188-218).
// The reference has no adjoint; this is the derivative of its own form.
//
// One kernel reads u, lam and G once and writes the gradient: the same
// element slab as axhelm (NP x NP threads per element, the t-column of u and
// lam in registers, the current layer in LDS), HBM-bound.
#include "element_kernels.h"

namespace nekamd {

template <int NP, bool G2>
__global__ void __launch_bounds__(NP * NP * epb<NP>())
    k_adj_grad(DMat<NP> Dm, const double *__restrict__ Dg, double *__restrict__ g1o,
               double *__restrict__ g2o, const double *__restrict__ u,
               const double *__restrict__ lam, const double *__restrict__ g, int64_t gstride,
               const double *__restrict__ bm1, int64_t nel) {
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, EPB = epb<NP>();
  __shared__ double s_D[NP][NP + 1];
  __shared__ double s_u[EPB][NP3];
  __shared__ double s_l[EPB][NP3];
  const int tid = threadIdx.x;
  const int le = tid / NP2;
  const int ij = tid - le * NP2;
  const int i = ij % NP, j = ij / NP;
  for (int t = tid; t < NP2; t += blockDim.x) s_D[t % NP][t / NP] = Dg[t];
  for (int64_t eb = blockIdx.x; eb * EPB < nel; eb += gridDim.x) {
    const int64_t e = eb * EPB + le;
    const bool valid = (le < EPB) && (e < nel);
    const int64_t base = e * NP3 + ij;
    double r_u[NP], r_l[NP];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double a = 0.0, b = 0.0;
      if (valid) {
        a = u[base + k * NP2];
        b = lam[base + k * NP2];
      }
      r_u[k] = a;
      r_l[k] = b;
      s_u[le][k * NP2 + ij] = a;
      s_l[le][k * NP2 + ij] = b;
    }
    __syncthreads();
    if (!valid) continue;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double ur = 0.0, us = 0.0, ut = 0.0, lr = 0.0, ls = 0.0, lt = 0.0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        const double dim = s_D[i][m], djm = s_D[j][m], dkm = Dm.d[k + NP * m];
        ur += dim * s_u[le][k * NP2 + j * NP + m];
        us += djm * s_u[le][k * NP2 + m * NP + i];
        ut += dkm * r_u[m];
        lr += dim * s_l[le][k * NP2 + j * NP + m];
        ls += djm * s_l[le][k * NP2 + m * NP + i];
        lt += dkm * r_l[m];
      }
      const int64_t l = base + k * NP2;
      const double G1 = g[l], G2v = g[gstride + l], G3 = g[2 * gstride + l];
      const double G4 = g[3 * gstride + l], G5 = g[4 * gstride + l], G6 = g[5 * gstride + l];
      const double wr = ur * G1 + us * G4 + ut * G5;
      const double ws = ur * G4 + us * G2v + ut * G6;
      const double wt = ur * G5 + us * G6 + ut * G3;
      g1o[l] = -(lr * wr + ls * ws + lt * wt);
      if constexpr (G2) g2o[l] = -(r_l[k] * bm1[l] * r_u[k]);
    }
  }
}

template <int NP>
static int adj_grad_dispatch(const Mesh &m, double *g1, double *g2, const double *u,
                             const double *lam, hipStream_t s) {
  DMat<NP> Dm;
  for (int t = 0; t < NP * NP; ++t) Dm.d[t] = m.D[t];
  constexpr int EPB = epb<NP>();
  const int64_t ngroups = (m.nel + EPB - 1) / EPB;
  int grid = (int)((ngroups < (int64_t)1 << 20) ? ngroups : ((int64_t)1 << 20));
  if (grid < 1) grid = 1;
  const dim3 block(NP * NP * EPB);
  if (g2)
    hipLaunchKernelGGL((k_adj_grad<NP, true>), dim3(grid), block, 0, s, Dm, m.Dd, g1, g2, u, lam,
                       m.g, m.n, m.bm1, m.nel);
  else
    hipLaunchKernelGGL((k_adj_grad<NP, false>), dim3(grid), block, 0, s, Dm, m.Dd, g1, g2, u,
                       lam, m.g, m.n, m.bm1, m.nel);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int launch_adj_grad(const Mesh &m, double *g1, double *g2, const double *u, const double *lam,
                    hipStream_t s) {
#define CALL(NP) adj_grad_dispatch<NP>(m, g1, g2, u, lam, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

}  // namespace nekamd
