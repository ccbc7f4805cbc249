// Mixed-precision pressure solve: hmh_flex_cg (hmholtz.f:2164-2281) with the
// h1mg V-cycle, to the FP64 tolerance, by defect correction.  The outer loop
// keeps the solution and the true residual b - A x in FP64 (axhelm,
// gather-scatter and norms in FP64); each correction is an inner flexible CG
// in single precision: its vectors, the operator (geometric factors stored
// in FP32) and the fine level of the V-cycle in FP32, its inner products
// accumulated in FP64; the local solves of the intermediate levels in FP32
// between FP64 residuals and corrections; the coarse solve in FP64, to a
// loose relative tolerance inside the inner V-cycles.  The
// inner solve stops at kInnerRtol times the outer residual (or at half the
// outer tolerance), so two or three outer steps reach 1e-8.
#include <algorithm>
#include <cmath>
#include <vector>

#include "element_kernels.h"
#include "hsmg.h"

namespace nekamd {

namespace {

constexpr int kMxBlocks = 2048;
constexpr double kInnerRtol = 1e-4;
constexpr int kMaxOuter = 8;

struct MxState {
  double rho0, rho1, rho2, beta, alpha, den, rnorm, tol;
  int iter, done;
};

#define NK_LOOP(i, n)                                                          \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ void part1(double a, double *part) {
  __shared__ double sa[4];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
  __syncthreads();
}

__device__ double fsum(const double *part, int nb) {
  __shared__ double sf[4];
  double v = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) v += part[q];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) sf[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = ((sf[0] + sf[1]) + sf[2]) + sf[3];
  __syncthreads();
  return t;
}

__global__ void k_mx_to_f32(float *__restrict__ d, const double *__restrict__ s, int64_t n) {
  NK_LOOP(i, n) d[i] = (float)s[i];
}

__global__ void k_mx_zero_flag(int *flag) { *flag = 0; }

// flag = 1 where a != b
__global__ void k_mx_neq(const double *a, const double *b, int64_t n, int *flag) {
  NK_LOOP(i, n)
  if (a[i] != b[i]) atomicOr(flag, 1);
}

template <int NP>
struct DMatF {
  float d[NP * NP];  // D(i,j) at d[i + NP*j]
};

// inner start: r32 = r64, x32 = 0, p32 = 0
__global__ void k_mx_start(float *__restrict__ r, float *__restrict__ x, float *__restrict__ p,
                           const double *__restrict__ r64, int64_t n) {
  NK_LOOP(i, n) {
    r[i] = (float)r64[i];
    x[i] = 0.0f;
    p[i] = 0.0f;
  }
}

__global__ void k_mx_state(MxState *st, double tol) {
  st->rho1 = 1.0;
  st->iter = 0;
  st->done = 0;
  st->tol = tol;
}

// the FP32 multiplicity weight of node i: the FP32 copy, or the count byte
// c decoded as the FP32 rounding of 1/c (what the copy holds when wt = 1/c)
__device__ __forceinline__ float wt32(const float *__restrict__ w, int64_t i) { return w[i]; }
__device__ __forceinline__ float wt32(const uint8_t *__restrict__ c, int64_t i) {
  return (float)(1.0 / (double)c[i]);
}

// partials of sum z*wt*r and sum z*wt*w in FP64 (rho1 and the rho2 term,
// as the FP64 solver: rho2 = -alpha sum z*wt*w_prev)
template <typename WT>
__global__ void __launch_bounds__(256)
    k_mx_dots(const float *__restrict__ z, const float *__restrict__ r,
              const float *__restrict__ w, const WT *__restrict__ wt, int64_t n, double *pa,
              double *pb) {
  double a = 0.0, b = 0.0;
  NK_LOOP(i, n) {
    const double zw = (double)z[i] * (double)wt32(wt, i);
    a += zw * (double)r[i];
    b += zw * (double)w[i];
  }
  part1(a, pa);
  part1(b, pb);
}

__global__ void __launch_bounds__(256) k_mx_beta(MxState *st, const double *pa, const double *pb) {
  const double a = fsum(pa, kMxBlocks);
  const double b = fsum(pb, kMxBlocks);
  if (threadIdx.x == 0) {
    st->rho0 = st->rho1;
    st->rho1 = a;
    st->rho2 = (st->iter == 0) ? a : -st->alpha * b;
    st->beta = st->rho2 / st->rho0;
  }
}

// one block of 1024 threads (the fused axhelm leaves up to 65536 partials)
__global__ void __launch_bounds__(1024) k_mx_alpha(MxState *st, const double *part, int count) {
  const double t = block_sum_partials(part, count);
  if (threadIdx.x == 0) {
    st->den = t;
    st->alpha = st->rho1 / t;
  }
}

// x += alpha p; r -= alpha w; partial sum r*r*wt (FP64)
template <typename WT>
__global__ void __launch_bounds__(256)
    k_mx_update(float *__restrict__ x, float *__restrict__ r, const float *__restrict__ p,
                const float *__restrict__ w, const WT *__restrict__ wt, const MxState *st,
                int64_t n, double *part) {
  // 4 nodes per step with nontemporal hints (the vectors are separate
  // 256-byte aligned allocations), the tail node by node
  typedef float v4 __attribute__((ext_vector_type(4)));
  const float alpha = (float)st->alpha;
  double a = 0.0;
  const int64_t n4 = n >> 2;
  v4 *x4 = reinterpret_cast<v4 *>(x);
  v4 *r4 = reinterpret_cast<v4 *>(r);
  const v4 *p4 = reinterpret_cast<const v4 *>(p);
  const v4 *w4 = reinterpret_cast<const v4 *>(w);
  NK_LOOP(i, n4) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(x4 + i) +
                                    alpha * __builtin_nontemporal_load(p4 + i),
                                x4 + i);
    const v4 ri = __builtin_nontemporal_load(r4 + i) - alpha * __builtin_nontemporal_load(w4 + i);
    __builtin_nontemporal_store(ri, r4 + i);
    v4 t;
    if constexpr (std::is_same<WT, float>::value) {
      t = __builtin_nontemporal_load(reinterpret_cast<const v4 *>(wt) + i);
    } else {  // four count bytes
      typedef unsigned char u4 __attribute__((ext_vector_type(4)));
      const u4 c = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(wt) + i);
      t.x = (float)(1.0 / (double)c.x);
      t.y = (float)(1.0 / (double)c.y);
      t.z = (float)(1.0 / (double)c.z);
      t.w = (float)(1.0 / (double)c.w);
    }
    a += (double)ri.x * (double)ri.x * (double)t.x + (double)ri.y * (double)ri.y * (double)t.y +
         (double)ri.z * (double)ri.z * (double)t.z + (double)ri.w * (double)ri.w * (double)t.w;
  }
  for (int64_t i = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    const float ri = r[i] - alpha * w[i];
    r[i] = ri;
    a += (double)ri * (double)ri * (double)wt32(wt, i);
  }
  part1(a, part);
}

__global__ void __launch_bounds__(256) k_mx_rnorm(MxState *st, const double *part, double vol) {
  const double t = fsum(part, kMxBlocks);
  if (threadIdx.x == 0) {
    const double rn = sqrt(t / vol);
    st->rnorm = rn;
    st->iter += 1;
    if (rn < st->tol) st->done = 1;
  }
}

// x64 += x32
__global__ void k_mx_accum(double *__restrict__ x, const float *__restrict__ e, int64_t n) {
  NK_LOOP(i, n) x[i] = x[i] + (double)e[i];
}

// r = (b - mask*Ax) (with Ax dssum'd); partial r*r*wt
__global__ void __launch_bounds__(256)
    k_mx_resid(double *__restrict__ r, const double *__restrict__ b,
               const double *__restrict__ ax, const double *__restrict__ mask,
               const double *__restrict__ wt, int64_t n, double *part) {
  double a = 0.0;
  NK_LOOP(i, n) {
    const double ri = b[i] - mask[i] * ax[i];
    r[i] = ri;
    a += ri * ri * wt[i];
  }
  part1(a, part);
}

// r = b * mgmask; partial r*r*wt
__global__ void __launch_bounds__(256)
    k_mx_init(double *__restrict__ r, double *__restrict__ x, const double *__restrict__ b,
              const double *__restrict__ mgmask, const double *__restrict__ wt, int64_t n,
              double *part) {
  double a = 0.0;
  NK_LOOP(i, n) {
    const double ri = b[i] * mgmask[i];
    r[i] = ri;
    x[i] = 0.0;
    a += ri * ri * wt[i];
  }
  part1(a, part);
}

}  // namespace

// Single-precision axhelm of the inner Krylov loops: p = beta*p + z
// (flexible CG) or, JAC, p = d*r + beta*p (Jacobi CG, z = r; returns while
// *donep), beta read in FP64 and rounded; w = A_L p with the FP32 factors
// (h1, h2 uniform), partial sums of p.w in FP64 per block.  Same element
// slab as the FP64 kernel.
__device__ __forceinline__ float ldnt32(const float *q) { return __builtin_nontemporal_load(q); }

template <int NP, bool H2, bool JAC>
__global__ void __launch_bounds__(NP * NP * epb<NP>())
    k_axhelm32(DMatF<NP> Dk, const double *__restrict__ Dg, float *__restrict__ w,
               float *__restrict__ p, const float *__restrict__ z, const float *__restrict__ dd,
               const float *__restrict__ g, int64_t gstride, float h1, float h2,
               const float *__restrict__ bm1, int64_t nel, const double *__restrict__ betap,
               const int *__restrict__ donep, double *__restrict__ den_part) {
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, EPB = epb<NP>();
  __shared__ float s_D[NP][NP + 1];
  __shared__ float s_u[EPB][NP3];
  __shared__ float s_wr[EPB][NP3];
  __shared__ float s_ws[EPB][NP3];
  __shared__ double s_red[NP2 * EPB];
  if constexpr (JAC) {
    if (*donep) return;
  }
  const float beta = (float)*betap;
  const int tid = threadIdx.x;
  const int le = tid / NP2;
  const int ij = tid - le * NP2;
  const int i = ij % NP, j = ij / NP;
  for (int t = tid; t < NP2; t += blockDim.x) s_D[t % NP][t / NP] = (float)Dg[t];
  double den = 0.0;
  for (int64_t eb = blockIdx.x; eb * EPB < nel; eb += gridDim.x) {
    const int64_t e = eb * EPB + le;
    const bool valid = (le < EPB) && (e < nel);
    const int64_t base = e * NP3 + ij;
    float r_u[NP], r_au[NP];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      float v = 0.0f;
      if (valid) {
        const int64_t l = base + k * NP2;
        if constexpr (JAC) v = ldnt32(dd + l) * ldnt32(z + l) + beta * ldnt32(p + l);
        else v = beta * ldnt32(p + l) + ldnt32(z + l);
        __builtin_nontemporal_store(v, p + l);
      }
      r_u[k] = v;
      s_u[le][k * NP2 + ij] = v;
      r_au[k] = 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      float ur = 0.0f, us = 0.0f, ut = 0.0f;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ur += s_D[i][m] * s_u[le][k * NP2 + j * NP + m];
        us += s_D[j][m] * s_u[le][k * NP2 + m * NP + i];
        ut += Dk.d[k + NP * m] * r_u[m];
      }
      float wr = 0.0f, ws = 0.0f, wt = 0.0f;
      if (valid) {
        const int64_t l = base + k * NP2;
        const float G1 = ldnt32(g + l), G2 = ldnt32(g + gstride + l),
                    G3 = ldnt32(g + 2 * gstride + l);
        const float G4 = ldnt32(g + 3 * gstride + l), G5 = ldnt32(g + 4 * gstride + l),
                    G6 = ldnt32(g + 5 * gstride + l);
        wr = h1 * (ur * G1 + us * G4 + ut * G5);
        ws = h1 * (ur * G4 + us * G2 + ut * G6);
        wt = h1 * (ur * G5 + us * G6 + ut * G3);
      }
      s_wr[le][k * NP2 + ij] = wr;
      s_ws[le][k * NP2 + ij] = ws;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dk.d[k + NP * kk] * wt;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      float ar = 0.0f, as = 0.0f;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ar += s_D[m][i] * s_wr[le][k * NP2 + j * NP + m];
        as += s_D[m][j] * s_ws[le][k * NP2 + m * NP + i];
      }
      if (valid) {
        const int64_t l = base + k * NP2;
        float a = ar + as + r_au[k];
        if constexpr (H2) a = a + (h2 * ldnt32(bm1 + l)) * r_u[k];
        __builtin_nontemporal_store(a, w + l);
        den += (double)a * (double)r_u[k];
      }
    }
  }
  const double t = block_sum(den, s_red);
  if (threadIdx.x == 0) den_part[blockIdx.x] = t;
}

// The same single-precision operator with the r- and s-direction
// contractions on the matrix cores (v_mfma_f32_16x16x4_f32, NP = 8..16), G
// t-planes staged at a time (G = the block's wavefronts), as k_axhelm_mxp
// does in FP64: per plane the gradients ur, us and the transposed sums
// ar + as are 16 x 16 products (zero-padded beyond NP) whose [j][i] tiles
// go one plane per wavefront; the owner thread of each t-column does the
// t-direction terms, the factors and the epilogue as in k_axhelm32.  Lane l
// of a product holds A[l%16][4c + l/16] and B[4c + l/16][l%16] of chunk c and
// gets D[4(l/16)+v][l%16] in register v (the FP32 layout, checked on
// gfx950).  Results agree with k_axhelm32 to FP32 rounding.
typedef float nk_f4 __attribute__((ext_vector_type(4)));

template <int NP, bool H2, bool JAC>
__global__ void __launch_bounds__(mx_threads<NP>())
    k_axhelm32_mxp(DMatF<NP> Dk, const double *__restrict__ Dg, float *__restrict__ w,
                   float *__restrict__ p, const float *__restrict__ z,
                   const float *__restrict__ dd, const float *__restrict__ g, int64_t gstride,
                   float h1, float h2, const float *__restrict__ bm1, int64_t nel,
                   const double *__restrict__ betap, const int *__restrict__ donep,
                   double *__restrict__ den_part) {
  static_assert(has_mfma<NP>(), "one 16 x 16 tile per plane");
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, T = mx_threads<NP>(), NW = T / 64, G = NW;
  constexpr int PJ = NP | 1, PP = PJ * NP, KC = (NP + 3) / 4, NG = (NP + G - 1) / G;
  __shared__ float s_u[G * PP];   // the group's planes of p, then ar + as
  __shared__ float s_wr[G * PP];  // ur, then wr
  __shared__ float s_ws[G * PP];  // us, then ws
  __shared__ double s_red[T];
  if constexpr (JAC) {
    if (*donep) return;
  }
  const float beta = (float)*betap;
  const int tid = threadIdx.x;
  const bool own = tid < NP2;
  const int i = own ? tid % NP : 0, j = own ? tid / NP : 0;
  const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lg = lane >> 4;
  const bool lin = lr < NP;
  float Dn[KC], Dt[KC];  // D(lr, 4c + lg) and D(4c + lg, lr), 0 outside NP x NP
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const int m = 4 * c + lg;
    const bool in = lin && m < NP;
    Dn[c] = in ? (float)Dg[lr + NP * m] : 0.0f;
    Dt[c] = in ? (float)Dg[m + NP * lr] : 0.0f;
  }
  const int own_at = i + PJ * j;
  float *const wu = s_u + wv * PP, *const wwr = s_wr + wv * PP, *const wws = s_ws + wv * PP;
  double den = 0.0;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    const int64_t base = e * NP3 + tid;
    float r_u[NP], r_au[NP], r_o[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      float v = 0.0f;
      if (own) {
        const int64_t l = base + k * NP2;
        if constexpr (JAC) v = ldnt32(dd + l) * ldnt32(z + l) + beta * ldnt32(p + l);
        else v = beta * ldnt32(p + l) + ldnt32(z + l);
        __builtin_nontemporal_store(v, p + l);
      }
      r_u[k] = v;
      r_au[k] = 0.0f;
      r_o[k] = 0.0f;
    }
#pragma unroll
    for (int gq = 0; gq < NG; ++gq) {
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // the previous group's (or element's) LDS reads are done
      if (own) {
#pragma unroll
        for (int q = 0; q < G; ++q)
          if (gq * G + q < NP) s_u[q * PP + own_at] = r_u[gq * G + q];
      }
      __syncthreads();
      const bool act = gq * G + wv < NP;  // wave-uniform
      if (act) {
        nk_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f}, acs = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const float a = (lin && m < NP) ? wu[PJ * lr + m] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Dn[c], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const float b = (lin && m < NP) ? wu[PJ * m + lr] : 0.0f;
          acs = __builtin_amdgcn_mfma_f32_16x16x4f32(Dn[c], b, acs, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int jj = 4 * lg + v;
          if (lin && jj < NP) {
            wwr[PJ * jj + lr] = acc[v];
            wws[PJ * jj + lr] = acs[v];
          }
        }
      }
      __syncthreads();
      if (own) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
          const int k = gq * G + q;
          if (k < NP) {
            const int at = q * PP + own_at;
            const float ur = s_wr[at], us = s_ws[at];
            float ut = 0.0f;
#pragma unroll
            for (int m = 0; m < NP; ++m) ut += Dk.d[k + NP * m] * r_u[m];
            const int64_t l = base + k * NP2;
            const float G1 = ldnt32(g + l), G2 = ldnt32(g + gstride + l),
                        G3 = ldnt32(g + 2 * gstride + l);
            const float G4 = ldnt32(g + 3 * gstride + l), G5 = ldnt32(g + 4 * gstride + l),
                        G6 = ldnt32(g + 5 * gstride + l);
            const float wr = h1 * (ur * G1 + us * G4 + ut * G5);
            const float ws = h1 * (ur * G4 + us * G2 + ut * G6);
            const float wt = h1 * (ur * G5 + us * G6 + ut * G3);
            s_wr[at] = wr;
            s_ws[at] = ws;
#pragma unroll
            for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dk.d[k + NP * kk] * wt;
          }
        }
      }
      __syncthreads();
      if (act) {
        nk_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const float a = (lin && m < NP) ? wwr[PJ * lr + m] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Dt[c], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const float b = (lin && m < NP) ? wws[PJ * m + lr] : 0.0f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(Dt[c], b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int jj = 4 * lg + v;
          if (lin && jj < NP) wu[PJ * jj + lr] = acc[v];
        }
      }
      __syncthreads();
      if (own) {
#pragma unroll
        for (int q = 0; q < G; ++q)
          if (gq * G + q < NP) r_o[gq * G + q] = s_u[q * PP + own_at];
      }
    }
    if (own) {
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int64_t l = base + k * NP2;
        float a = r_o[k] + r_au[k];
        if constexpr (H2) a = a + (h2 * ldnt32(bm1 + l)) * r_u[k];
        __builtin_nontemporal_store(a, w + l);
        den += (double)a * (double)r_u[k];
      }
    }
  }
  const double t = block_sum(den, s_red);
  if (threadIdx.x == 0) den_part[blockIdx.x] = t;
}

// Kernel of the FP32 operator per order: 0 vector FMA (k_axhelm32), 1 the
// matrix cores (k_axhelm32_mxp, lx1 = 8..16); nkh_axhelm32_set_variant fixes
// it.  Defaults from 100 mixed Jacobi-PCG iterations per kernel on boxes of
// about 5e7 unique DOF (MI355X, GDOF*iter/s vector -> matrix cores): lx1 = 8
// 30.9 -> 31.6 (and the mixed flexible CG at E = 64^3 122.9-123.3 -> 122.1-
// 122.8 ms per solve), 11 30.3 -> 32.0, 14 26.2 -> 27.3, 15 25.4 -> 28.1;
// the other orders are faster with vector FMA or within 1 % (lx1 = 12:
// 32.3 -> 29.0).
int g_ax32_variant[kMaxNP + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0};

template <int NP, bool JAC>
static int axhelm32_dispatch(const Mesh &m, float *w, float *p, const float *z, const float *dd,
                             float h1, float h2, const double *betap, const int *donep,
                             double *den_part, hipStream_t s) {
  DMatF<NP> Dm;
  for (int t = 0; t < NP * NP; ++t) Dm.d[t] = (float)m.D[t];
  const int grid = axhelm_cg_grid(m);
  if constexpr (has_mfma<NP>()) {
    if (g_ax32_variant[NP] == 1) {
      const dim3 bm(mx_threads<NP>());
      if (h2 != 0.0f)
        hipLaunchKernelGGL((k_axhelm32_mxp<NP, true, JAC>), dim3(grid), bm, 0, s, Dm, m.Dd, w, p,
                           z, dd, m.g32, m.n, h1, h2, m.bm1_32, m.nel, betap, donep, den_part);
      else
        hipLaunchKernelGGL((k_axhelm32_mxp<NP, false, JAC>), dim3(grid), bm, 0, s, Dm, m.Dd, w,
                           p, z, dd, m.g32, m.n, h1, h2, m.bm1_32, m.nel, betap, donep,
                           den_part);
      NK_HIP(hipGetLastError());
      return NKH_OK;
    }
  }
  const dim3 block(NP * NP * epb<NP>());
  if (h2 != 0.0f)
    hipLaunchKernelGGL((k_axhelm32<NP, true, JAC>), dim3(grid), block, 0, s, Dm, m.Dd, w, p, z,
                       dd, m.g32, m.n, h1, h2, m.bm1_32, m.nel, betap, donep, den_part);
  else
    hipLaunchKernelGGL((k_axhelm32<NP, false, JAC>), dim3(grid), block, 0, s, Dm, m.Dd, w, p, z,
                       dd, m.g32, m.n, h1, h2, m.bm1_32, m.nel, betap, donep, den_part);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int launch_axhelm32_jac(const Mesh &m, float *w, float *p, const float *r, const float *d,
                        float h1, float h2, const CgState *st, double *rho_part,
                        hipStream_t s) {
#define CALL(NP) \
  axhelm32_dispatch<NP, true>(m, w, p, r, d, h1, h2, &st->beta, &st->done, rho_part, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

__global__ void k_to_f32(float *__restrict__ d, const double *__restrict__ s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}

int axhelm32_set_variant(int np, int v) {
  NK_CHECK(np >= 2 && np <= kMaxNP && (v == 0 || v == 1), NKH_ERR_ARG,
           "FP32 axhelm variant: lx1=%d variant=%d (0 or 1)", np, v);
  NK_CHECK(v == 0 || (np >= 8 && np <= 16), NKH_ERR_UNSUPPORTED,
           "FP32 axhelm on the matrix cores: lx1=%d outside 8..16", np);
  g_ax32_variant[np] = v;
  return NKH_OK;
}

int axhelm32_variant(int np, int *v) {
  NK_CHECK(np >= 2 && np <= kMaxNP, NKH_ERR_ARG, "FP32 axhelm variant: lx1=%d", np);
  *v = g_ax32_variant[np];
  return NKH_OK;
}

int ensure_geom32(Mesh &m, hipStream_t s) {
  const int64_t n = m.n;
  if (!m.g32) {
    if (hipMalloc((void **)&m.g32, sizeof(float) * 6 * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&m.bm1_32, sizeof(float) * (size_t)n) != hipSuccess) {
      (void)hipGetLastError();
      if (m.g32) (void)hipFree(m.g32);
      m.g32 = nullptr;
      m.bm1_32 = nullptr;
      set_error("FP32 operator: out of device memory (%lld floats)", (long long)(7 * n));
      return NKH_ERR_NOMEM;
    }
    m.g32_version = -1;
  }
  if (m.g32_version != m.geom_version) {
    const int64_t g6 = 6 * n;
    const int grid6 = (int)std::min<int64_t>((g6 + 255) / 256, 65536);
    const int grid1 = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_f32, dim3(grid6), dim3(256), 0, s, m.g32, m.g, g6);
    hipLaunchKernelGGL(k_to_f32, dim3(grid1), dim3(256), 0, s, m.bm1_32, m.bm1, n);
    NK_HIP(hipGetLastError());
    m.g32_version = m.geom_version;
  }
  return NKH_OK;
}

namespace {

int falloc(float **p, int64_t n) {
  if (hipMalloc((void **)p, sizeof(float) * (size_t)n) != hipSuccess) {
    (void)hipGetLastError();
    set_error("mixed solver: out of device memory (%lld floats)", (long long)n);
    return NKH_ERR_NOMEM;
  }
  return NKH_OK;
}

int launch_axhelm32_flex(const Mesh &m, float *w, float *p, const float *z, float h1, float h2,
                         const MxState *st, double *den_part, hipStream_t s) {
#define CALL(NP) \
  axhelm32_dispatch<NP, false>(m, w, p, z, nullptr, h1, h2, &st->beta, nullptr, den_part, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

}  // namespace

int run_flex_cg_mixed(Mesh &m, double *res, Coef h1, Coef h2, const double *wt,
                      const double *mask, int maxit, const nkh_params &prm, nkh_result *out,
                      double *rnorm_history) {
  NK_CHECK(m.mg, NKH_ERR_STATE, "hmh_flex_cg needs the multigrid hierarchy (nkh_mg_setup)");
  NK_CHECK(prm.param40 == 0 || prm.param40 == 1, NKH_ERR_UNSUPPORTED,
           "hmh_flex_cg: coarse solver param40=%d (0 exact, 1 one AMG V-cycle)", prm.param40);
  NK_CHECK(h1.uniform() && h2.uniform(), NKH_ERR_UNSUPPORTED,
           "mixed-precision solve: h1 and h2 must be uniform");
  Mg &mg = *m.mg;
  hipStream_t s = cur_stream();
  const int64_t n = m.n;
  const double *mgmask = mg.lev[mg.lmax].mask;
  // the inner solver needs the identities of the fused FP64 path: pmask
  // equal to the multigrid's fine mask, ids zero at element interiors
  bool fused = false;
  NK_TRY(fused_conditions(m, mask, wt, nullptr, &fused, s));
  NK_CHECK(fused, NKH_ERR_UNSUPPORTED,
           "mixed-precision solve: needs 0/1 masks equal at shared nodes, mult summing to 1 and "
           "ids zero at element interiors");
  {
    int differ = 0;
    if (!mg.fflag) NK_HIP(hipMalloc(&mg.fflag, sizeof(int)));
    hipLaunchKernelGGL(k_mx_zero_flag, dim3(1), dim3(1), 0, s, mg.fflag);
    hipLaunchKernelGGL(k_mx_neq, dim3(mg_grid(n)), dim3(256), 0, s, mask, mgmask, n, mg.fflag);
    NK_HIP(hipMemcpyAsync(&differ, mg.fflag, sizeof(int), hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    NK_CHECK(!differ, NKH_ERR_UNSUPPORTED,
             "mixed-precision solve: the pressure mask must be the multigrid's fine mask");
  }
  if (!mg.xr) {
    for (float **p : {&mg.xr, &mg.xp, &mg.xz, &mg.xw, &mg.xx, &mg.xwt}) NK_TRY(falloc(p, n));
    NK_TRY(mg_dalloc(&mg.xpart, 3 * kMxBlocks + kAxBlocks));
    NK_HIP(hipMalloc(&mg.xst, sizeof(MxState)));
  }
  if (!mg.fr)
    for (double **p : {&mg.fr, &mg.fp, &mg.fz, &mg.fw}) NK_TRY(mg_dalloc(p, n));
  const dim3 G(kMxBlocks), B(256);
  NK_TRY(ensure_geom32(m, s));
  hipLaunchKernelGGL(k_mx_to_f32, dim3(mg_grid(n)), B, 0, s, mg.xwt, wt, n);
  // the weights as count bytes where every one is 1/count (1 B instead of 4
  // per node in the inner dots and update)
  const uint8_t *wtc = nullptr;
  NK_TRY(mg_wt_codes(mg, wt, n, s, &wtc));
  MxState *st = (MxState *)mg.xst;
  double *pa = mg.xpart, *pb = mg.xpart + kMxBlocks, *pc = mg.xpart + 2 * kMxBlocks;
  double *den_part = mg.xpart + 3 * kMxBlocks;
  const int nden = axhelm_cg_grid(m);
  double *r64 = mg.fr, *x64 = mg.fp, *b64 = mg.fz, *ax = mg.fw;
  // tolerance as hmh_flex_cg (hmholtz.f:2205-2220)
  double tolps = 0.0;
  NK_TRY(run_chktcg1(m, &tolps, res, h1, h2, mask, wt));
  if (prm.param21 > 0 && tolps > std::fabs(prm.param21)) tolps = std::fabs(prm.param21);
  if (prm.istep == 0) tolps = 1.e-4;
  NK_HIP(hipMemcpyAsync(b64, res, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_mx_init, G, B, 0, s, r64, x64, b64, mgmask, wt, n, pc);
  auto host_sum = [&](double *part, double *v) -> int {
    std::vector<double> h(kMxBlocks);
    NK_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * kMxBlocks, hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
    double a = 0.0;
    for (double x : h) a += x;
    *v = a;
    return NKH_OK;
  };
  double t0;
  NK_TRY(host_sum(pc, &t0));
  const double div0 = std::sqrt(t0 / m.vol);
  const double tolpss = prm.param21 < 0 ? std::fabs(prm.param21) * div0 : tolps;
  double rn = div0;
  const float h1f = (float)h1.s, h2f = (float)h2.s;
  int total = 0;
  MxState hs{};
  for (int outer = 0; outer < kMaxOuter && rn >= tolpss && total < maxit; ++outer) {
    const double itol = std::max(kInnerRtol * rn, 0.5 * tolpss);
    hipLaunchKernelGGL(k_mx_start, dim3(mg_grid(n)), B, 0, s, mg.xr, mg.xx, mg.xp, r64, n);
    hipLaunchKernelGGL(k_mx_state, dim3(1), dim3(1), 0, s, st, itol);
    NK_HIP(hipGetLastError());
    while (total < maxit) {
      // the inner V-cycles' coarse problem: one AMG V-cycle for param40 = 1,
      // else the coarse AMG-PCG to 0.1 relative (crs_mode 2)
      NK_TRY(mg_solve32(m, mg.xz, mg.xr, s, prm.param40 == 1 ? 1 : 2));
      if (wtc)
        hipLaunchKernelGGL(k_mx_dots<uint8_t>, G, B, 0, s, mg.xz, mg.xr, mg.xw, wtc, n, pa, pb);
      else
        hipLaunchKernelGGL(k_mx_dots<float>, G, B, 0, s, mg.xz, mg.xr, mg.xw, mg.xwt, n, pa, pb);
      hipLaunchKernelGGL(k_mx_beta, dim3(1), B, 0, s, st, pa, pb);
      NK_TRY(launch_axhelm32_flex(m, mg.xw, mg.xp, mg.xz, h1f, h2f, st, den_part, s));
      NK_TRY(gs_assemble(m.gs, mg.xw, nullptr, s));
      hipLaunchKernelGGL(k_mx_alpha, dim3(1), dim3(1024), 0, s, st, den_part, nden);
      if (wtc)
        hipLaunchKernelGGL(k_mx_update<uint8_t>, G, B, 0, s, mg.xx, mg.xr, mg.xp, mg.xw, wtc, st,
                           n, pa);
      else
        hipLaunchKernelGGL(k_mx_update<float>, G, B, 0, s, mg.xx, mg.xr, mg.xp, mg.xw, mg.xwt,
                           st, n, pa);
      hipLaunchKernelGGL(k_mx_rnorm, dim3(1), B, 0, s, st, pa, m.vol);
      NK_HIP(hipGetLastError());
      NK_HIP(hipMemcpyAsync(&hs, st, sizeof(MxState), hipMemcpyDeviceToHost, s));
      NK_HIP(hipStreamSynchronize(s));
      if (rnorm_history) rnorm_history[total] = hs.rnorm;
      total += 1;
      if (hs.done) break;
    }
    // x += e; r = b - A x in FP64 (the true residual)
    hipLaunchKernelGGL(k_mx_accum, dim3(mg_grid(n)), B, 0, s, x64, mg.xx, n);
    NK_TRY(launch_axhelm(m, ax, x64, h1, h2, s));
    NK_TRY(gs_op(m.gs, ax, 0, s));
    hipLaunchKernelGGL(k_mx_resid, G, B, 0, s, r64, b64, ax, mask, wt, n, pc);
    NK_HIP(hipGetLastError());
    double t;
    NK_TRY(host_sum(pc, &t));
    rn = std::sqrt(t / m.vol);
  }
  NK_HIP(hipMemcpyAsync(res, x64, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  // hmholtz.f:2272: orthogonalise the solution against the null space
  if (mg_null_space_on(mg, h2)) NK_TRY(mg_ortho(mg, res, n, s));
  NK_HIP(hipStreamSynchronize(s));
  if (out) {
    out->niter = total;
    out->converged = rn < tolpss;
    out->rbn0 = div0;
    out->rbn2 = rn;
    out->tol = tolpss;
  }
  return NKH_OK;
}

}  // namespace nekamd
