// Element kernels of the Helmholtz path: axhelm (plain and fused into the
// CG iteration), geometric factors, Jacobi diagonal -- templates shared by
// the translation units that instantiate them.
//
// Work decomposition (all kernels): one NP x NP thread slab per element, the
// slab walks the NP layers of the element in registers (the t-direction
// column of u lives in the thread's registers, the current r/s layer in LDS).
// At NP = 8 the slab is exactly one 64-lane wavefront.  Elements are visited
// with a grid-stride loop over a fixed grid so partial sums land in a fixed
// number of slots.
#pragma once
#include "nekamd_internal.h"

namespace nekamd {

template <int NP>
struct DMat {
  double d[NP * NP];  // D(i,j) at d[i + NP*j]  (dxm1, column-major)
};

// Elements per block: pack small elements so a block is at least 64 lanes.
template <int NP>
constexpr int epb() {
  return (NP * NP >= 64) ? 1 : (64 / (NP * NP));
}

// The wide packing (the autotuner's variants 2 and 3): the number of
// NP x NP slabs per block, at most 512 threads, that leaves the fewest idle
// lanes in the block's last wavefront (the smallest such count on ties).
// At NP = 5 two slabs fill 50 of 64 lanes, five fill 125 of 128; at NP = 12
// one fills 144 of 192, three fill 432 of 448.
template <int NP>
constexpr int epb_wide() {
  int best = epb<NP>();
  double waste = 1.0;
  for (int e = 1; e * NP * NP <= 512; ++e) {
    const int t = e * NP * NP, w = (t + 63) / 64 * 64;
    const double f = double(w - t) / double(w);
    if (f < waste - 1e-12) {
      waste = f;
      best = e;
    }
  }
  return best;
}
template <int NP>
constexpr bool has_wide() {
  return epb_wide<NP>() != epb<NP>();
}
// the element-staged kernel with the wide packing only where its LDS
// (3 NP^3 doubles per slab) stays within 96 KB, so two blocks share a CU
template <int NP>
constexpr bool has_wide_elem() {
  return has_wide<NP>() && 3 * NP * NP * NP * 8 * epb_wide<NP>() <= 96 * 1024;
}

// Sum over the block's threads through LDS (blockDim need not fill whole
// wavefronts), in a fixed tree order -> deterministic.  sbuf holds blockDim
// doubles.  Result valid on thread 0.
__device__ __forceinline__ double block_sum(double v, double *sbuf) {
  const int t = threadIdx.x, nt = blockDim.x;
  __syncthreads();
  sbuf[t] = v;
  __syncthreads();
  for (int h = 1; h < nt; h <<= 1) {
    if ((t % (2 * h)) == 0 && t + h < nt) sbuf[t] += sbuf[t + h];
    __syncthreads();
  }
  return sbuf[0];
}

// Streaming accesses of the element kernels: every field word is read or
// written once per apply, so loads and stores carry the nontemporal hint
// (measured on MI355X: +9-17 % on a plain copy).
__device__ __forceinline__ double ldnt(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(double v, double *p) { __builtin_nontemporal_store(v, p); }

// ---------------------------------------------------------------------------
// axhelm: hmholtz.f:188-218 (3-D deformed branch) and :225 (h2 term).
//   tmp1 = h1*(g1*ur + g4*us + g5*ut)   (same order as col3/addcol3 there)
//   tmp2 = h1*(g2*us + g4*ur + g6*ut)
//   tmp3 = h1*(g3*ut + g5*ur + g6*us)
//   au   = D^T_r tmp1 + D^T_s tmp2 + D^T_t tmp3  (+ (h2*bm1)*u)
// MODE 0: au = A u.  MODE 1 (CG): p = d*r + beta*p is formed on the fly and
// stored, w = A_L p (unassembled), and sum_l p_l w_l -- the CG denominator
// rho = (Q Q^T w, p)_mult for a continuous, masked p -- is written to one
// slot per block when rho_part is given.  MODE 2 (flexible CG): the same
// with p = beta*p + z (add2s1, hmholtz.f:2243), z passed as r, beta read
// from *betap; with xacc, also the previous iteration's solution update
// res += alpha p (hmholtz.f:2250-2251) on the p it reads, alpha from
// *alphap -- the same arithmetic one iteration later, so that the update
// kernel need not read p and x again.  MODE 3: MODE 0 gated by a state (returns while st->done),
// for applies that a captured iteration runs only on some passes.
// ---------------------------------------------------------------------------
template <int NP, int MODE, bool H1F, int H2M, int EPB_ = epb<NP>()>
__global__ void __launch_bounds__(NP * NP * EPB_)
    k_axhelm(DMat<NP> Dm, const double *__restrict__ Dg, double *__restrict__ au,
             const double *__restrict__ uin, const double *__restrict__ g, int64_t gstride,
             const double *__restrict__ h1f, double h1s, const double *__restrict__ h2f,
             double h2s, const double *__restrict__ bm1, int64_t nel,
             double *__restrict__ pout, const double *__restrict__ r,
             const double *__restrict__ dinv, const double *__restrict__ mult,
             const CgState *__restrict__ st, double *__restrict__ rho_part,
             const double *__restrict__ betap, double *__restrict__ xacc,
             const double *__restrict__ alphap) {
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, EPB = EPB_;
  __shared__ double s_D[NP][NP + 1];
  __shared__ double s_u[EPB][NP3];
  __shared__ double s_wr[EPB][NP3];
  __shared__ double s_ws[EPB][NP3];
  __shared__ double s_red[NP2 * EPB];

  double beta = 0.0;
  if constexpr (MODE == 1) {
    if (st->done) return;
    beta = st->beta;
  }
  double alpha = 0.0;  // MODE 2 with xacc: x += alpha p (the previous p)
  if constexpr (MODE == 2) {
    beta = *betap;
    if (xacc) alpha = *alphap;
  }
  if constexpr (MODE == 3) {
    if (st->done) return;
  }
  const int tid = threadIdx.x;
  const int le = tid / NP2;
  const int ij = tid - le * NP2;
  const int i = ij % NP, j = ij / NP;
  for (int t = tid; t < NP2; t += blockDim.x) s_D[t % NP][t / NP] = Dg[t];

  double rho = 0.0;
  for (int64_t eb = blockIdx.x; eb * EPB < nel; eb += gridDim.x) {
    const int64_t e = eb * EPB + le;
    const bool valid = (le < EPB) && (e < nel);
    const int64_t base = e * NP3 + ij;
    double r_u[NP], r_au[NP];
    __syncthreads();  // LDS reuse across element groups (and s_D fill)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double v = 0.0;
      if (valid) {
        const int64_t l = base + k * NP2;
        if constexpr (MODE == 1) {
          // add2s1: p = beta*p + z, z = r*d
          v = beta * ldnt(uin + l) + ldnt(dinv + l) * ldnt(r + l);
          stnt(v, pout + l);
        } else if constexpr (MODE == 2) {
          const double po = ldnt(uin + l);
          v = beta * po + ldnt(r + l);
          if (xacc) stnt(ldnt(xacc + l) + alpha * po, xacc + l);
          stnt(v, pout + l);
        } else {
          v = ldnt(uin + l);
        }
      }
      r_u[k] = v;
      s_u[le][k * NP2 + ij] = v;
      r_au[k] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double ur = 0.0, us = 0.0, ut = 0.0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ur += s_D[i][m] * s_u[le][k * NP2 + j * NP + m];
        us += s_D[j][m] * s_u[le][k * NP2 + m * NP + i];
        ut += Dm.d[k + NP * m] * r_u[m];
      }
      double wr = 0.0, ws = 0.0, wt = 0.0;
      if (valid) {
        const int64_t l = base + k * NP2;
        const double G1 = ldnt(g + l), G2 = ldnt(g + gstride + l),
                     G3 = ldnt(g + 2 * gstride + l);
        const double G4 = ldnt(g + 3 * gstride + l), G5 = ldnt(g + 4 * gstride + l),
                     G6 = ldnt(g + 5 * gstride + l);
        wr = ur * G1;
        ws = us * G2;
        wt = ut * G3;
        wr = wr + us * G4;
        wr = wr + ut * G5;
        ws = ws + ur * G4;
        ws = ws + ut * G6;
        wt = wt + ur * G5;
        wt = wt + us * G6;
        const double h1 = H1F ? ldnt(h1f + l) : h1s;
        wr *= h1;
        ws *= h1;
        wt *= h1;
      }
      s_wr[le][k * NP2 + ij] = wr;
      s_ws[le][k * NP2 + ij] = ws;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dm.d[k + NP * kk] * wt;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double ar = 0.0, as = 0.0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ar += s_D[m][i] * s_wr[le][k * NP2 + j * NP + m];
        as += s_D[m][j] * s_ws[le][k * NP2 + m * NP + i];
      }
      if (valid) {
        const int64_t l = base + k * NP2;
        double a = ar + as + r_au[k];
        if constexpr (H2M == 1) a = a + (h2s * ldnt(bm1 + l)) * r_u[k];
        if constexpr (H2M == 2) a = a + (ldnt(h2f + l) * ldnt(bm1 + l)) * r_u[k];
        stnt(a, au + l);
        if constexpr (MODE == 1 || MODE == 2) rho += a * r_u[k];  // p . (A_L p), element-local
      }
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    if (rho_part != nullptr) {
      const double t = block_sum(rho, s_red);
      if (threadIdx.x == 0) rho_part[blockIdx.x] = t;
    }
  }
}

// The same operator with the element staged in LDS one t-plane at a time:
// 3 NP^2 doubles of LDS per element instead of 3 NP^3 (at NP = 16 the
// element-staged kernel holds 96 KB and one block per CU).  ur, us, wr and
// ws of plane k need only plane k; the t-direction terms stay in registers.
// Arithmetic and its order are those of k_axhelm, so results are identical.
template <int NP, int MODE, bool H1F, int H2M, int EPB_ = epb<NP>()>
__global__ void __launch_bounds__(NP * NP * EPB_)
    k_axhelm_pl(DMat<NP> Dm, const double *__restrict__ Dg, double *__restrict__ au,
                const double *__restrict__ uin, const double *__restrict__ g, int64_t gstride,
                const double *__restrict__ h1f, double h1s, const double *__restrict__ h2f,
                double h2s, const double *__restrict__ bm1, int64_t nel,
                double *__restrict__ pout, const double *__restrict__ r,
                const double *__restrict__ dinv, const double *__restrict__ mult,
                const CgState *__restrict__ st, double *__restrict__ rho_part,
                const double *__restrict__ betap, double *__restrict__ xacc,
             const double *__restrict__ alphap) {
  constexpr int NP2 = NP * NP, EPB = EPB_;
  __shared__ double s_D[NP][NP + 1];
  __shared__ double s_u[EPB][NP2];
  __shared__ double s_wr[EPB][NP2];
  __shared__ double s_ws[EPB][NP2];
  __shared__ double s_red[NP2 * EPB];
  double beta = 0.0;
  if constexpr (MODE == 1) {
    if (st->done) return;
    beta = st->beta;
  }
  double alpha = 0.0;  // MODE 2 with xacc: x += alpha p (the previous p)
  if constexpr (MODE == 2) {
    beta = *betap;
    if (xacc) alpha = *alphap;
  }
  if constexpr (MODE == 3) {
    if (st->done) return;
  }
  const int tid = threadIdx.x;
  const int le = tid / NP2;
  const int ij = tid - le * NP2;
  const int i = ij % NP, j = ij / NP;
  for (int t = tid; t < NP2; t += blockDim.x) s_D[t % NP][t / NP] = Dg[t];
  double rho = 0.0;
  for (int64_t eb = blockIdx.x; eb * EPB < nel; eb += gridDim.x) {
    const int64_t e = eb * EPB + le;
    const bool valid = (le < EPB) && (e < nel);
    const int64_t base = e * (int64_t)(NP2 * NP) + ij;
    double r_u[NP], r_au[NP], r_o[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double v = 0.0;
      if (valid) {
        const int64_t l = base + k * NP2;
        if constexpr (MODE == 1) {
          v = beta * ldnt(uin + l) + ldnt(dinv + l) * ldnt(r + l);
          stnt(v, pout + l);
        } else if constexpr (MODE == 2) {
          const double po = ldnt(uin + l);
          v = beta * po + ldnt(r + l);
          if (xacc) stnt(ldnt(xacc + l) + alpha * po, xacc + l);
          stnt(v, pout + l);
        } else {
          v = ldnt(uin + l);
        }
      }
      r_u[k] = v;
      r_au[k] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      // one plane at a time: keeps the scheduler from hoisting the next
      // planes' geometric-factor loads (6 doubles each) into registers
      __builtin_amdgcn_sched_barrier(0);
      __syncthreads();  // the previous plane (or element) is read
      s_u[le][ij] = r_u[k];
      __syncthreads();
      double ur = 0.0, us = 0.0, ut = 0.0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ur += s_D[i][m] * s_u[le][j * NP + m];
        us += s_D[j][m] * s_u[le][m * NP + i];
        ut += Dm.d[k + NP * m] * r_u[m];
      }
      double wr = 0.0, ws = 0.0, wt = 0.0;
      if (valid) {
        const int64_t l = base + k * NP2;
        const double G1 = ldnt(g + l), G2 = ldnt(g + gstride + l),
                     G3 = ldnt(g + 2 * gstride + l);
        const double G4 = ldnt(g + 3 * gstride + l), G5 = ldnt(g + 4 * gstride + l),
                     G6 = ldnt(g + 5 * gstride + l);
        wr = ur * G1;
        ws = us * G2;
        wt = ut * G3;
        wr = wr + us * G4;
        wr = wr + ut * G5;
        ws = ws + ur * G4;
        ws = ws + ut * G6;
        wt = wt + ur * G5;
        wt = wt + us * G6;
        const double h1 = H1F ? ldnt(h1f + l) : h1s;
        wr *= h1;
        ws *= h1;
        wt *= h1;
      }
      s_wr[le][ij] = wr;
      s_ws[le][ij] = ws;
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dm.d[k + NP * kk] * wt;
      __syncthreads();
      double ar = 0.0, as = 0.0;
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        ar += s_D[m][i] * s_wr[le][j * NP + m];
        as += s_D[m][j] * s_ws[le][m * NP + i];
      }
      r_o[k] = ar + as;
      // pin this plane's sums here: the compiler would otherwise sink every
      // plane's contractions to the final stores and keep their LDS operands
      // (2 NP doubles per plane) live in registers
      asm volatile("" : "+v"(r_o[k]));
#pragma unroll
      for (int kk = 0; kk < NP; ++kk) asm volatile("" : "+v"(r_au[kk]));
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int64_t l = base + k * NP2;
      double a = r_o[k] + r_au[k];
      if constexpr (H2M == 1) a = a + (h2s * (valid ? ldnt(bm1 + l) : 0.0)) * r_u[k];
      if constexpr (H2M == 2)
        a = a + ((valid ? ldnt(h2f + l) : 0.0) * (valid ? ldnt(bm1 + l) : 0.0)) * r_u[k];
      if (valid) stnt(a, au + l);
      if constexpr (MODE == 1 || MODE == 2) rho += a * r_u[k];  // a = r_u = 0 off the mesh
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    if (rho_part != nullptr) {
      const double t = block_sum(rho, s_red);
      if (threadIdx.x == 0) rho_part[blockIdx.x] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// The same operator with the r- and s-direction contractions on the matrix
// cores (v_mfma_f64_16x16x4_f64), for NP = 8..16 (the autotuner's variant 4).
// One element per block of whole wavefronts; thread (i, j) < NP^2 owns the
// t-column as in k_axhelm.  Per t-plane k the contractions are 16 x 16
// products (zero-padded beyond NP, K in chunks of 4) that give [j][i] tiles:
//   ur^T_k[j][i] = sum_m u(m,j,k) D(i,m)     us^T_k[j][i] = sum_m D(j,m) u(i,m,k)
//   (ar + as)^T_k[j][i] = sum_m wr(m,j,k) D(m,i) + sum_m D(m,j) ws(i,m,k)
// the 2 NP gradient products and the NP output products dealt over the
// block's wavefronts.  Lane l of a product holds A[l%16][4c + l/16] and
// B[4c + l/16][l%16] of chunk c and gets D[4v + l/16][l%16] in register v
// (the FP64 layout, checked on gfx950).  The D fragments stay in registers;
// u, ur/wr and us/ws are staged in LDS with an odd j-pitch, so the 16 rows of
// a fragment fall in different banks.  The t-direction terms, the geometric
// factors and the epilogues are those of k_axhelm, in the same order; the
// matrix cores sum each product in their own order, so au agrees with the
// other kernels to rounding, not bit for bit.
// ---------------------------------------------------------------------------
typedef double nk_d4 __attribute__((ext_vector_type(4)));

template <int NP>
constexpr bool has_mfma() {
  return NP >= 8 && NP <= 16;
}
template <int NP>
constexpr int mx_threads() {
  return (NP * NP + 63) / 64 * 64;
}

template <int NP, int MODE, bool H1F, int H2M>
__global__ void __launch_bounds__(mx_threads<NP>())
    k_axhelm_mx(DMat<NP> Dm, const double *__restrict__ Dg, double *__restrict__ au,
                const double *__restrict__ uin, const double *__restrict__ g, int64_t gstride,
                const double *__restrict__ h1f, double h1s, const double *__restrict__ h2f,
                double h2s, const double *__restrict__ bm1, int64_t nel,
                double *__restrict__ pout, const double *__restrict__ r,
                const double *__restrict__ dinv, const double *__restrict__ mult,
                const CgState *__restrict__ st, double *__restrict__ rho_part,
                const double *__restrict__ betap, double *__restrict__ xacc,
             const double *__restrict__ alphap) {
  static_assert(has_mfma<NP>(), "one 16 x 16 tile per plane");
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, NW = mx_threads<NP>() / 64;
  constexpr int PJ = NP | 1, PK = PJ * NP, KC = (NP + 3) / 4;
  __shared__ double s_u[PK * NP];   // u, then ar + as
  __shared__ double s_wr[PK * NP];  // ur, then wr
  __shared__ double s_ws[PK * NP];  // us, then ws
  double beta = 0.0;
  if constexpr (MODE == 1) {
    if (st->done) return;
    beta = st->beta;
  }
  double alpha = 0.0;  // MODE 2 with xacc: x += alpha p (the previous p)
  if constexpr (MODE == 2) {
    beta = *betap;
    if (xacc) alpha = *alphap;
  }
  if constexpr (MODE == 3) {
    if (st->done) return;
  }
  const int tid = threadIdx.x;
  const bool own = tid < NP2;
  const int i = own ? tid % NP : 0, j = own ? tid / NP : 0;
  const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lg = lane >> 4;
  const bool lin = lr < NP;
  // Dn[c] = D(lr, 4c + lg), Dt[c] = D(4c + lg, lr); 0 outside NP x NP
  double Dn[KC], Dt[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const int m = 4 * c + lg;
    const bool in = lin && m < NP;
    Dn[c] = in ? Dg[lr + NP * m] : 0.0;
    Dt[c] = in ? Dg[m + NP * lr] : 0.0;
  }
  const int own_at = i + PJ * j;  // the owner's node in a plane of LDS
  double rho = 0.0;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    const int64_t base = e * NP3 + tid;
    double r_u[NP], r_au[NP];
    __syncthreads();  // the previous element's LDS reads are done
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double v = 0.0;
      if (own) {
        const int64_t l = base + k * NP2;
        if constexpr (MODE == 1) {
          v = beta * ldnt(uin + l) + ldnt(dinv + l) * ldnt(r + l);
          stnt(v, pout + l);
        } else if constexpr (MODE == 2) {
          const double po = ldnt(uin + l);
          v = beta * po + ldnt(r + l);
          if (xacc) stnt(ldnt(xacc + l) + alpha * po, xacc + l);
          stnt(v, pout + l);
        } else {
          v = ldnt(uin + l);
        }
        s_u[k * PK + own_at] = v;
      }
      r_u[k] = v;
      r_au[k] = 0.0;
    }
    __syncthreads();
    // ur and us of every plane: product p = 2k + (0: ur, 1: us)
    for (int p = wv; p < 2 * NP; p += NW) {
      const int k = p >> 1;
      const double *uk = s_u + k * PK;
      nk_d4 acc = {0.0, 0.0, 0.0, 0.0};
      if ((p & 1) == 0) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double a = (lin && m < NP) ? uk[PJ * lr + m] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Dn[c], acc, 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double b = (lin && m < NP) ? uk[PJ * m + lr] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Dn[c], b, acc, 0, 0, 0);
        }
      }
      double *dst = ((p & 1) == 0 ? s_wr : s_ws) + k * PK;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jj = 4 * v + lg;
        if (lin && jj < NP) dst[PJ * jj + lr] = acc[v];
      }
    }
    __syncthreads();
    if (own) {
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int at = k * PK + own_at;
        const double ur = s_wr[at], us = s_ws[at];
        double ut = 0.0;
#pragma unroll
        for (int m = 0; m < NP; ++m) ut += Dm.d[k + NP * m] * r_u[m];
        const int64_t l = base + k * NP2;
        const double G1 = ldnt(g + l), G2 = ldnt(g + gstride + l),
                     G3 = ldnt(g + 2 * gstride + l);
        const double G4 = ldnt(g + 3 * gstride + l), G5 = ldnt(g + 4 * gstride + l),
                     G6 = ldnt(g + 5 * gstride + l);
        double wr = ur * G1;
        double ws = us * G2;
        double wt = ut * G3;
        wr = wr + us * G4;
        wr = wr + ut * G5;
        ws = ws + ur * G4;
        ws = ws + ut * G6;
        wt = wt + ur * G5;
        wt = wt + us * G6;
        const double h1 = H1F ? ldnt(h1f + l) : h1s;
        wr *= h1;
        ws *= h1;
        wt *= h1;
        s_wr[at] = wr;
        s_ws[at] = ws;
#pragma unroll
        for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dm.d[k + NP * kk] * wt;
      }
    }
    __syncthreads();
    // ar + as of every plane into s_u (u itself is in the owners' registers)
    for (int k = wv; k < NP; k += NW) {
      const double *wrk = s_wr + k * PK, *wsk = s_ws + k * PK;
      nk_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const int m = 4 * c + lg;
        const double a = (lin && m < NP) ? wrk[PJ * lr + m] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Dt[c], acc, 0, 0, 0);
      }
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const int m = 4 * c + lg;
        const double b = (lin && m < NP) ? wsk[PJ * m + lr] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Dt[c], b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jj = 4 * v + lg;
        if (lin && jj < NP) s_u[k * PK + PJ * jj + lr] = acc[v];
      }
    }
    __syncthreads();
    if (own) {
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int64_t l = base + k * NP2;
        double a = s_u[k * PK + own_at] + r_au[k];
        if constexpr (H2M == 1) a = a + (h2s * ldnt(bm1 + l)) * r_u[k];
        if constexpr (H2M == 2) a = a + (ldnt(h2f + l) * ldnt(bm1 + l)) * r_u[k];
        stnt(a, au + l);
        if constexpr (MODE == 1 || MODE == 2) rho += a * r_u[k];
      }
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    if (rho_part != nullptr) {
      const double t = block_sum(rho, s_u);  // after the last element's reads of s_u
      if (threadIdx.x == 0) rho_part[blockIdx.x] = t;
    }
  }
}

// The matrix-core operator staged G t-planes at a time (G = the block's
// wavefronts; the autotuner's variant 5): 3 G NP^2 doubles of LDS instead of
// 3 NP^3 (NP = 16: 26 KB, not 104), so several blocks share a CU.  Per group
// of planes each wavefront forms ur and us of its plane, the owners do the
// pointwise work of the group's planes (t-direction terms in registers as in
// k_axhelm_pl), each wavefront forms ar + as of its plane, and the owners
// keep those in registers until the t-direction sums are complete.  The
// products, their layouts and the arithmetic order are those of
// k_axhelm_mx, so the two give the same au bit for bit.
template <int NP, int MODE, bool H1F, int H2M>
__global__ void __launch_bounds__(mx_threads<NP>())
    k_axhelm_mxp(DMat<NP> Dm, const double *__restrict__ Dg, double *__restrict__ au,
                 const double *__restrict__ uin, const double *__restrict__ g, int64_t gstride,
                 const double *__restrict__ h1f, double h1s, const double *__restrict__ h2f,
                 double h2s, const double *__restrict__ bm1, int64_t nel,
                 double *__restrict__ pout, const double *__restrict__ r,
                 const double *__restrict__ dinv, const double *__restrict__ mult,
                 const CgState *__restrict__ st, double *__restrict__ rho_part,
                 const double *__restrict__ betap, double *__restrict__ xacc,
             const double *__restrict__ alphap) {
  static_assert(has_mfma<NP>(), "one 16 x 16 tile per plane");
  constexpr int NP2 = NP * NP, NP3 = NP2 * NP, NW = mx_threads<NP>() / 64, G = NW;
  constexpr int PJ = NP | 1, PP = PJ * NP, KC = (NP + 3) / 4, NG = (NP + G - 1) / G;
  __shared__ double s_u[G * PP];   // the group's planes of u, then ar + as
  __shared__ double s_wr[G * PP];  // ur, then wr
  __shared__ double s_ws[G * PP];  // us, then ws
  double beta = 0.0;
  if constexpr (MODE == 1) {
    if (st->done) return;
    beta = st->beta;
  }
  double alpha = 0.0;  // MODE 2 with xacc: x += alpha p (the previous p)
  if constexpr (MODE == 2) {
    beta = *betap;
    if (xacc) alpha = *alphap;
  }
  if constexpr (MODE == 3) {
    if (st->done) return;
  }
  const int tid = threadIdx.x;
  const bool own = tid < NP2;
  const int i = own ? tid % NP : 0, j = own ? tid / NP : 0;
  const int lane = tid & 63, wv = tid >> 6, lr = lane & 15, lg = lane >> 4;
  const bool lin = lr < NP;
  double Dn[KC], Dt[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const int m = 4 * c + lg;
    const bool in = lin && m < NP;
    Dn[c] = in ? Dg[lr + NP * m] : 0.0;
    Dt[c] = in ? Dg[m + NP * lr] : 0.0;
  }
  const int own_at = i + PJ * j;
  double *const wu = s_u + wv * PP, *const wwr = s_wr + wv * PP, *const wws = s_ws + wv * PP;
  double rho = 0.0;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    const int64_t base = e * NP3 + tid;
    double r_u[NP], r_au[NP], r_o[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      double v = 0.0;
      if (own) {
        const int64_t l = base + k * NP2;
        if constexpr (MODE == 1) {
          v = beta * ldnt(uin + l) + ldnt(dinv + l) * ldnt(r + l);
          stnt(v, pout + l);
        } else if constexpr (MODE == 2) {
          const double po = ldnt(uin + l);
          v = beta * po + ldnt(r + l);
          if (xacc) stnt(ldnt(xacc + l) + alpha * po, xacc + l);
          stnt(v, pout + l);
        } else {
          v = ldnt(uin + l);
        }
      }
      r_u[k] = v;
      r_au[k] = 0.0;
      r_o[k] = 0.0;
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
      const bool act = gq * G + wv < NP;  // wave-uniform: this wave's plane exists
      if (act) {
        nk_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double a = (lin && m < NP) ? wu[PJ * lr + m] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Dn[c], acc, 0, 0, 0);
        }
        nk_d4 acs = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double b = (lin && m < NP) ? wu[PJ * m + lr] : 0.0;
          acs = __builtin_amdgcn_mfma_f64_16x16x4f64(Dn[c], b, acs, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int jj = 4 * v + lg;
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
            const double ur = s_wr[at], us = s_ws[at];
            double ut = 0.0;
#pragma unroll
            for (int m = 0; m < NP; ++m) ut += Dm.d[k + NP * m] * r_u[m];
            const int64_t l = base + k * NP2;
            const double G1 = ldnt(g + l), G2 = ldnt(g + gstride + l),
                         G3 = ldnt(g + 2 * gstride + l);
            const double G4 = ldnt(g + 3 * gstride + l), G5 = ldnt(g + 4 * gstride + l),
                         G6 = ldnt(g + 5 * gstride + l);
            double wr = ur * G1;
            double ws = us * G2;
            double wt = ut * G3;
            wr = wr + us * G4;
            wr = wr + ut * G5;
            ws = ws + ur * G4;
            ws = ws + ut * G6;
            wt = wt + ur * G5;
            wt = wt + us * G6;
            const double h1 = H1F ? ldnt(h1f + l) : h1s;
            wr *= h1;
            ws *= h1;
            wt *= h1;
            s_wr[at] = wr;
            s_ws[at] = ws;
#pragma unroll
            for (int kk = 0; kk < NP; ++kk) r_au[kk] += Dm.d[k + NP * kk] * wt;
          }
        }
      }
      __syncthreads();
      if (act) {
        nk_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double a = (lin && m < NP) ? wwr[PJ * lr + m] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Dt[c], acc, 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int m = 4 * c + lg;
          const double b = (lin && m < NP) ? wws[PJ * m + lr] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Dt[c], b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int jj = 4 * v + lg;
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
        double a = r_o[k] + r_au[k];
        if constexpr (H2M == 1) a = a + (h2s * ldnt(bm1 + l)) * r_u[k];
        if constexpr (H2M == 2) a = a + (ldnt(h2f + l) * ldnt(bm1 + l)) * r_u[k];
        stnt(a, au + l);
        if constexpr (MODE == 1 || MODE == 2) rho += a * r_u[k];
      }
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    if (rho_part != nullptr) {
      const double t = block_sum(rho, s_u);  // after the last group's reads of s_u
      if (threadIdx.x == 0) rho_part[blockIdx.x] = t;
    }
  }
}

// Per polynomial order: which of the two kernels (0: element staged in LDS,
// 1: t-planes) for the plain apply (and its gated form), the CG-fused and
// the flexible-CG-fused apply, and how many blocks for the plain apply;
// chosen by the autotuner on the first large apply of that order.
struct AxhelmChoice {
  int variant = 0;       // plain apply (MODE 0, MODE 3)
  int grid = 0;          // plain apply: blocks (0 -> one per element group, up to 2^20)
  int variant_cg = 0;    // MODE 1
  int variant_flex = 0;  // MODE 2
  bool tuned = false;
};
inline AxhelmChoice g_axhelm_choice[kMaxNP + 1];

// One launch of the axhelm kernel (element- or plane-staged, EPB slabs per
// block; MX 1 or 2: the matrix-core kernel staging the element or groups of
// planes, one element per block) for the apply MODE.
template <int NP, int MODE, bool PLANE, int EPB, int MX = 0>
static int axhelm_go(const Mesh &m, const DMat<NP> &Dm, double *au, const double *u, Coef h1,
                     Coef h2, double *pout, const double *r, const double *dinv,
                     const double *mult, const CgState *st, double *rho_part, hipStream_t s,
                     const double *betap, int gridcap, double *xacc = nullptr,
                     const double *alphap = nullptr) {
  const int64_t ngroups = (m.nel + EPB - 1) / EPB;
  int grid;
  if (MODE == 3) {
    // gated: most launches find the gate closed, and an empty launch costs
    // in proportion to its blocks: 16 element groups per block, 64 to 4096
    int64_t g = ngroups / 16;
    g = g < 64 ? 64 : (g > 4096 ? 4096 : g);
    grid = (int)(ngroups < g ? ngroups : g);
  } else if (MODE >= 1) {
    // one partial sum per block: the fixed count the CG's scalar kernels read
    grid = axhelm_cg_grid(m);
  } else {
    const int64_t cap = gridcap > 0 ? gridcap : ((int64_t)1 << 20);
    grid = (int)(ngroups < cap ? ngroups : cap);
  }
  if (grid < 1) grid = 1;
  const dim3 block(MX != 0 ? mx_threads<NP>() : NP * NP * EPB);
  // h2 mode: 0 none (reference ifh2 false), 1 uniform, 2 field
  const int h2m = h2.uniform() ? (h2.s != 0.0 ? 1 : 0) : 2;
  const bool h1f = !h1.uniform();
#define NK_LAUNCH(H1F_, H2M_)                                                                \
  if constexpr (MX == 1)                                                                      \
    hipLaunchKernelGGL((k_axhelm_mx<NP, MODE, H1F_, H2M_>), dim3(grid), block, 0, s, Dm,      \
                       m.Dd, au, u, m.g, m.n, h1.f, h1.s, h2.f, h2.s, m.bm1, m.nel, pout, r,   \
                       dinv, mult, st, rho_part, betap, xacc, alphap);                        \
  else if constexpr (MX == 2)                                                                 \
    hipLaunchKernelGGL((k_axhelm_mxp<NP, MODE, H1F_, H2M_>), dim3(grid), block, 0, s, Dm,     \
                       m.Dd, au, u, m.g, m.n, h1.f, h1.s, h2.f, h2.s, m.bm1, m.nel, pout, r,   \
                       dinv, mult, st, rho_part, betap, xacc, alphap);                        \
  else if constexpr (PLANE)                                                                   \
    hipLaunchKernelGGL((k_axhelm_pl<NP, MODE, H1F_, H2M_, EPB>), dim3(grid), block, 0, s, Dm, \
                       m.Dd, au, u, m.g, m.n, h1.f, h1.s, h2.f, h2.s, m.bm1, m.nel, pout, r,   \
                       dinv, mult, st, rho_part, betap, xacc, alphap);                        \
  else                                                                                        \
    hipLaunchKernelGGL((k_axhelm<NP, MODE, H1F_, H2M_, EPB>), dim3(grid), block, 0, s, Dm,    \
                       m.Dd, au, u, m.g, m.n, h1.f, h1.s, h2.f, h2.s, m.bm1, m.nel, pout, r,   \
                       dinv, mult, st, rho_part, betap, xacc, alphap)
  if (h1f) {
    if (h2m == 0) NK_LAUNCH(true, 0);
    else if (h2m == 1) NK_LAUNCH(true, 1);
    else NK_LAUNCH(true, 2);
  } else {
    if (h2m == 0) NK_LAUNCH(false, 0);
    else if (h2m == 1) NK_LAUNCH(false, 1);
    else NK_LAUNCH(false, 2);
  }
#undef NK_LAUNCH
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// Variants: 0 element-staged, 1 plane-staged (epb slabs per block), 2
// plane-staged and 3 element-staged with the wide packing (epb_wide), 4 and
// 5 the r/s contractions on the matrix cores (NP = 8..16) with the element
// or groups of planes staged; an order without a distinct wide packing runs
// 1 for 2 and 0 for 3, as does 3 where the wide element-staged kernel would
// not fit two blocks per CU, and an order outside 8..16 runs 0 for 4 and 5.
// The plain apply is element-local, so variants 0-3 give the same au bit
// for bit, as do 4 and 5; the fused applies' partial sums of p.w group the
// elements by block, so the wide packings round those sums differently;
// the matrix cores agree with the vector kernels to rounding.
template <int NP, int MODE>
static int axhelm_dispatch(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                           double *pout, const double *r, const double *dinv,
                           const double *mult, const CgState *st, double *rho_part,
                           hipStream_t s, const double *betap = nullptr,
                           double *xacc = nullptr, const double *alphap = nullptr) {
  DMat<NP> Dm;
  for (int t = 0; t < NP * NP; ++t) Dm.d[t] = m.D[t];
  const AxhelmChoice ch = g_axhelm_choice[NP];
  const int var = MODE == 1 ? ch.variant_cg : (MODE == 2 ? ch.variant_flex : ch.variant);
  const int gc = ch.grid;
#define NK_GO(PLANE_, EPB_)                                                                  \
  return axhelm_go<NP, MODE, PLANE_, EPB_>(m, Dm, au, u, h1, h2, pout, r, dinv, mult, st,   \
                                          rho_part, s, betap, gc, xacc, alphap)
  if (var == 2) {
    if constexpr (has_wide<NP>()) NK_GO(true, epb_wide<NP>());
    else NK_GO(true, epb<NP>());
  }
  if (var == 3) {
    if constexpr (has_wide_elem<NP>()) NK_GO(false, epb_wide<NP>());
    else NK_GO(false, epb<NP>());
  }
  if (var == 4 || var == 5) {
    if constexpr (has_mfma<NP>()) {
      if (var == 4)
        return axhelm_go<NP, MODE, false, 1, 1>(m, Dm, au, u, h1, h2, pout, r, dinv, mult, st,
                                                rho_part, s, betap, gc, xacc, alphap);
      return axhelm_go<NP, MODE, false, 1, 2>(m, Dm, au, u, h1, h2, pout, r, dinv, mult, st,
                                              rho_part, s, betap, gc, xacc, alphap);
    } else {
      NK_GO(false, epb<NP>());
    }
  }
  if (var == 1) NK_GO(true, epb<NP>());
  NK_GO(false, epb<NP>());
#undef NK_GO
}

// which variants an order has beyond 0 and 1 (bit 2: variant 2, bit 3: 3,
// bit 4: 4 and 5)
template <int NP>
constexpr int axhelm_wide_mask() {
  return (has_wide<NP>() ? 4 : 0) | (has_wide_elem<NP>() ? 8 : 0) | (has_mfma<NP>() ? 16 : 0);
}

#define NK_NP_SWITCH(np, CALL)                                      \
  switch (np) {                                                     \
    case 2: return CALL(2);                                         \
    case 3: return CALL(3);                                         \
    case 4: return CALL(4);                                         \
    case 5: return CALL(5);                                         \
    case 6: return CALL(6);                                         \
    case 7: return CALL(7);                                         \
    case 8: return CALL(8);                                         \
    case 9: return CALL(9);                                         \
    case 10: return CALL(10);                                       \
    case 11: return CALL(11);                                       \
    case 12: return CALL(12);                                       \
    case 13: return CALL(13);                                       \
    case 14: return CALL(14);                                       \
    case 15: return CALL(15);                                       \
    case 16: return CALL(16);                                       \
    default:                                                        \
      set_error("lx1=%d outside the supported range 2..16", np);    \
      return NKH_ERR_UNSUPPORTED;                                   \
  }

}  // namespace nekamd
