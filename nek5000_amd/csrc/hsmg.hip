// H1 multigrid V-cycle (hsmg.f:1855-1949 h1mg_solve) and the flexible CG
// around it (hmholtz.f:2164-2281 hmh_flex_cg), on the GPU.
//
// Schwarz smoother of level l (hsmg.f:425-494 h1mg_schwarz): the reference
// copies r into (n+2)^3 extended arrays and exchanges the ghost layers with a
// gather-scatter over an extended numbering, twice.  The exchange only ever
// moves one face layer between face neighbours, so here it is two gathers
// through maps built once from that numbering (hsmg_setup.hip):
//   k_sch_fdm : per element, the extended array is assembled in LDS from the
//               element's own r and the neighbours' first interior layers
//               (gin); the fast-diagonalisation solve S D S^T runs in LDS
//               and registers; the interior and the ghost faces of the local
//               solution are written out;
//   k_sch_out : e = interior solution + the neighbours' ghost faces that
//               land on this element's first interior layer (gout);
//   gs(level) : e summed at shared nodes.
// The weights (mask * Schwarz weight) are applied where the prolongation adds
// the coarser correction.  The bandwidth-bound head of the coarse chain
// (restrictions and the level-2 dssum) runs before the fine-level smoother;
// the coarse solve and the level-2 Schwarz then run on two more streams,
// beside it, and join before the last prolongation.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hsmg.h"

namespace nekamd {

namespace {

// odd LDS row pitch for the (n+2)^3 block, so that pencils along r are
// read without bank conflicts
// Launch shape of the local FDM solves: each thread owns PPT pencils of NL
// points per contraction (so each S value read from LDS feeds PPT FMAs); P
// is an odd LDS row pitch so that pencils along r are read without bank
// conflicts.
template <int NL>
struct Fdm {
  static constexpr int P = NL | 1;
  static constexpr int PPT = NL >= 8 ? 2 : 1;
  static constexpr int kActive = (NL * NL + PPT - 1) / PPT;
  static constexpr int kThreads = ((kActive + 63) / 64) * 64;
};

// Local Schwarz solves on level with nh = NL-2 points (hsmg.f:897-929
// hsmg_do_fast): w = (St (x) Ss (x) Sr) D (St^T (x) Ss^T (x) Sr^T) u.  u is
// the extended array: interior = r of the element, ghost faces = r of the
// face neighbours one layer inside them (0 where there is none), ghost edges
// and corners 0 -- what hsmg.f:455-467 leaves in the extended array.
// D(i,j,k) = 1/(lr_i + ls_j + lt_k) where that sum exceeds eps, else 0
// (hsmg.f:753-770), formed from the per-element eigenvalues.  One element
// per block.  A contraction along one direction gives each thread whole
// pencils along that direction in registers; the element's S and
// eigenvalues are staged in LDS with the input (one coalesced load phase)
// and read there as wave-uniform operands.  Output: the interior of w, and
// each ghost face value of w sent to the neighbour's slot it coincides with
// (gout), where the overlap sum reads it.
template <int NL, typename R, typename RI>
__global__ void __launch_bounds__(Fdm<NL>::kThreads)
    k_sch_fdm(R *__restrict__ wint, R *__restrict__ wgh, const RI *__restrict__ r,
              const int32_t *__restrict__ gin, const int32_t *__restrict__ gout,
              const R *__restrict__ S, const R *__restrict__ lam, int64_t nel) {
  constexpr int N = NL - 2, N2 = N * N, N3 = N2 * N, M2 = NL * NL, NLAM = 3 * NL + 1;
  constexpr int P = Fdm<NL>::P, PL = P * NL, T = Fdm<NL>::kThreads;
  constexpr int PPT = Fdm<NL>::PPT, NA = Fdm<NL>::kActive;
  __shared__ R u[PL * NL];
  __shared__ R sS[3 * M2];
  __shared__ R sl[NLAM];
  const int tid = threadIdx.x;
  // ghost face slot (f,a,b) -> LDS position of the extended array
  auto gpos = [](int q) {
    const int f = q / N2, a = q % N, b = (q / N) % N;
    int i, j, k;
    switch (f) {
      case 0: i = 0; j = a + 1; k = b + 1; break;
      case 1: i = NL - 1; j = a + 1; k = b + 1; break;
      case 2: i = a + 1; j = 0; k = b + 1; break;
      case 3: i = a + 1; j = NL - 1; k = b + 1; break;
      case 4: i = a + 1; j = b + 1; k = 0; break;
      default: i = a + 1; j = b + 1; k = NL - 1; break;
    }
    return i + P * j + PL * k;
  };
  // pencils of this thread: p = tid + q*NA (q < PPT), valid when p < M2;
  // (pa, pb) = the two coordinates that fix the pencil
  int pa[PPT], pb[PPT];
  bool pv[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int p = tid + q * NA;
    pv[q] = (tid < NA) && (p < M2);
    pa[q] = pv[q] ? p % NL : 0;
    pb[q] = pv[q] ? p / NL : 0;
  }
  // contraction of every pencil along direction D (0: r, 1: s, 2: t) with
  // the matrix Sm: y_i = sum_m Sm(m,i) x_m when TR, else Sm(i,m) x_m
  auto contract = [&](auto dir_tag, auto tr_tag, const R *Sm) {
    constexpr int D = decltype(dir_tag)::value;
    constexpr bool TR = decltype(tr_tag)::value;
    constexpr int st = D == 0 ? 1 : (D == 1 ? P : PL);
    R x[PPT][NL];
    int base[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      base[q] = D == 0 ? P * pa[q] + PL * pb[q] : (D == 1 ? pa[q] + PL * pb[q] : pa[q] + P * pb[q]);
      if (pv[q]) {
#pragma unroll
        for (int m = 0; m < NL; ++m) x[q][m] = u[base[q] + st * m];
      }
    }
#pragma unroll 2
    for (int i = 0; i < NL; ++i) {
      R acc[PPT];
#pragma unroll
      for (int q = 0; q < PPT; ++q) acc[q] = R(0);
#pragma unroll
      for (int m = 0; m < NL; ++m) {
        const R sv = TR ? Sm[m + NL * i] : Sm[i + NL * m];
#pragma unroll
        for (int q = 0; q < PPT; ++q) acc[q] += sv * x[q][m];
      }
#pragma unroll
      for (int q = 0; q < PPT; ++q)
        if (pv[q]) u[base[q] + st * i] = acc[q];
    }
  };
  using R_ = std::integral_constant<int, 0>;
  using S_ = std::integral_constant<int, 1>;
  using T_ = std::integral_constant<int, 2>;
  using TR_ = std::true_type;
  using NT_ = std::false_type;
  // element order: consecutive blocks go to different XCDs (round-robin), so
  // block b takes the elements of chunk b % 8, and face neighbours -- which
  // read each other's data -- are processed by one XCD around the same time
  // (its own L2).  ep -> e is a bijection of [0, nel) when gridDim.x is a
  // multiple of 8.
  const int64_t q8 = nel / 8, r8 = nel % 8;
  for (int64_t ep = blockIdx.x; ep < nel; ep += gridDim.x) {
    const int64_t x8 = ep % 8;
    const int64_t e = x8 * q8 + (x8 < r8 ? x8 : r8) + ep / 8;
    const RI *re = r + e * N3;
    const int32_t *ge = gin + e * 6 * N2;
    const R *Se = S + e * 3 * M2;
    // every load of the element issued before the first wait
    constexpr int NR = (N3 + T - 1) / T, NS = (3 * M2 + T - 1) / T, NG = (6 * N2 + T - 1) / T;
    R vr[NR], vs[NS], vg[NG], vl;
    int32_t gi[NG], gd[NG];
    const int32_t *go = gout + e * 6 * N2;
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T, qc = q < 6 * N2 ? q : 0;
      gi[it] = q < 6 * N2 ? ge[qc] : -1;
      gd[it] = q < 6 * N2 ? go[qc] : -1;
    }
#pragma unroll
    for (int it = 0; it < NR; ++it) {
      const int q = tid + it * T;
      vr[it] = R(re[q < N3 ? q : N3 - 1]);
    }
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * T;
      vs[it] = Se[q < 3 * M2 ? q : 3 * M2 - 1];
    }
    vl = lam[e * NLAM + (tid < NLAM ? tid : NLAM - 1)];
#pragma unroll
    for (int it = 0; it < NG; ++it) vg[it] = R(r[gi[it] >= 0 ? gi[it] : 0]);
    __syncthreads();  // the previous element's LDS reads are done
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * T;
      if (q < 3 * M2) sS[q] = vs[it];
    }
    if (tid < NLAM) sl[tid] = vl;
#pragma unroll
    for (int it = 0; it < NR; ++it) {
      const int q = tid + it * T;
      if (q < N3) {
        const int i = q % N, j = (q / N) % N, k = q / N2;
        u[(i + 1) + P * (j + 1) + PL * (k + 1)] = vr[it];
      }
    }
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T;
      if (q < 6 * N2) u[gpos(q)] = gi[it] >= 0 ? vg[it] : R(0);
    }
    for (int q = tid; q < M2 * NL; q += T) {  // ghost edges and corners
      const int i = q % NL, j = (q / NL) % NL, k = q / M2;
      const int g = (i == 0 || i == NL - 1) + (j == 0 || j == NL - 1) + (k == 0 || k == NL - 1);
      if (g >= 2) u[i + P * j + PL * k] = R(0);
    }
    __syncthreads();
    contract(R_(), TR_(), sS);
    __syncthreads();
    contract(S_(), TR_(), sS + M2);
    __syncthreads();
    contract(T_(), TR_(), sS + 2 * M2);
    // the diagonal, on the pencils just written by this thread
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (!pv[q]) continue;
      const R lij = sl[pa[q]] + sl[NL + pb[q]], eps = sl[3 * NL];
      const int base = pa[q] + P * pb[q];
#pragma unroll
      for (int k = 0; k < NL; ++k) {
        const R diag = lij + sl[2 * NL + k];
        const R v = u[base + PL * k];
        u[base + PL * k] = (diag > eps) ? v * (R(1) / diag) : R(0);
      }
    }
    contract(T_(), NT_(), sS + 2 * M2);
    __syncthreads();
    contract(S_(), NT_(), sS + M2);
    __syncthreads();
    contract(R_(), NT_(), sS);
    __syncthreads();
    R *wi = wint + e * N3;
#pragma unroll
    for (int it = 0; it < (N3 + T - 1) / T; ++it) {
      const int q = tid + it * T;
      if (q < N3) {
        const int i = q % N, j = (q / N) % N, k = q / N2;
        wi[q] = u[(i + 1) + P * (j + 1) + PL * (k + 1)];
      }
    }
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T;
      if (gd[it] >= 0) wgh[gd[it]] = u[gpos(q)];
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 local FDM solves (the drop-in's smoother): the same arithmetic as
// k_sch_fdm in the same order, organised for the FP64 vector units of
// CDNA4:
//   * the element's three NL x NL eigenvector matrices are read from LDS into
//     registers, 16 consecutive entries per register across the 16 lanes of
//     each row, and fed to the FMAs by a DPP row broadcast
//     (v_fmac_f64 ... row_newbcast:n): no LDS traffic for S inside the
//     contractions;
//   * every global load of an element (interior, ghost ids and the gathered
//     neighbour layers, S, eigenvalues, the ghost destinations) is issued
//     before the first one is waited for;
//   * the t contraction, the diagonal and the transposed t contraction run
//     in registers on each lane's t-pencils, without an LDS round trip;
//   * 1/(lr + ls + lt) by the hardware reciprocal refined with two
//     Newton-Raphson steps r <- r + r (1 - d r).
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ void fma_rowbcast(double &acc, double s, double x) {
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(acc)
      : "v"(s), "v"(x), "i"(L));
}

// y[q][i] = sum_m Sm(i,m) x[q][m] for the PPT pencils of a lane (TR: Sm(m,i)),
// Sm entry f = row + NL * col held in register f / 16 of lane f % 16 of each
// row; outputs two at a time (independent accumulator chains)
template <int NL, int PPT, bool TR, int I, int M>
__device__ __forceinline__ void fdm_fma2(double (&acc)[2][PPT], const double *sr,
                                         const double (&x)[PPT][NL]) {
  constexpr int f0 = TR ? M + NL * I : I + NL * M;
#pragma unroll
  for (int q = 0; q < PPT; ++q) fma_rowbcast<f0 % 16>(acc[0][q], sr[f0 / 16], x[q][M]);
  if constexpr (I + 1 < NL) {
    constexpr int f1 = TR ? M + NL * (I + 1) : (I + 1) + NL * M;
#pragma unroll
    for (int q = 0; q < PPT; ++q) fma_rowbcast<f1 % 16>(acc[1][q], sr[f1 / 16], x[q][M]);
  }
}
template <int NL, int PPT, bool TR, int I, int... Ms>
__device__ __forceinline__ void fdm_rows(double (&acc)[2][PPT], const double *sr,
                                         const double (&x)[PPT][NL],
                                         std::integer_sequence<int, Ms...>) {
  (fdm_fma2<NL, PPT, TR, I, Ms>(acc, sr, x), ...);
}
template <int NL, int PPT, bool TR, int... Ps>
__device__ __forceinline__ void fdm_matvec(double (&y)[PPT][NL], const double *sr,
                                           const double (&x)[PPT][NL],
                                           std::integer_sequence<int, Ps...>) {
  auto pair = [&](auto p_tag) {
    constexpr int I = 2 * decltype(p_tag)::value;
    double acc[2][PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) acc[0][q] = acc[1][q] = 0.0;
    fdm_rows<NL, PPT, TR, I>(acc, sr, x, std::make_integer_sequence<int, NL>());
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      y[q][I] = acc[0][q];
      if constexpr (I + 1 < NL) y[q][I + 1] = acc[1][q];
    }
  };
  (pair(std::integral_constant<int, Ps>()), ...);
}

template <int NL>
__global__ void __launch_bounds__(Fdm<NL>::kThreads)
    k_sch_fdm_d(double *__restrict__ wint, double *__restrict__ wgh,
                const double *__restrict__ r, const int32_t *__restrict__ gin,
                const int32_t *__restrict__ gout, const double *__restrict__ S,
                const double *__restrict__ lam, int64_t nel) {
  constexpr int N = NL - 2, N2 = N * N, N3 = N2 * N, M2 = NL * NL, NLAM = 3 * NL + 1;
  constexpr int P = Fdm<NL>::P, PL = P * NL, T = Fdm<NL>::kThreads;
  constexpr int PPT = Fdm<NL>::PPT, NA = Fdm<NL>::kActive;
  constexpr int NSR = (M2 + 15) / 16;  // registers per matrix
  constexpr int NR = (N3 + T - 1) / T, NS = (3 * M2 + T - 1) / T, NG = (6 * N2 + T - 1) / T;
  __shared__ double u[PL * NL];
  __shared__ double sS[3 * M2 + 16];
  __shared__ double sl[NLAM];
  const int tid = threadIdx.x;
  auto gpos = [](int q) {
    const int f = q / N2, a = q % N, b = (q / N) % N;
    int i, j, k;
    switch (f) {
      case 0: i = 0; j = a + 1; k = b + 1; break;
      case 1: i = NL - 1; j = a + 1; k = b + 1; break;
      case 2: i = a + 1; j = 0; k = b + 1; break;
      case 3: i = a + 1; j = NL - 1; k = b + 1; break;
      case 4: i = a + 1; j = b + 1; k = 0; break;
      default: i = a + 1; j = b + 1; k = NL - 1; break;
    }
    return i + P * j + PL * k;
  };
  // pencils of this thread: p = tid + q*NA (q < PPT), valid when p < M2;
  // an invalid pencil reads pencil 0 and its results are dropped (all lanes
  // take part in the broadcast FMAs)
  int pa[PPT], pb[PPT];
  bool pv[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int p = tid + q * NA;
    pv[q] = (tid < NA) && (p < M2);
    pa[q] = pv[q] ? p % NL : 0;
    pb[q] = pv[q] ? p / NL : 0;
  }
  double sr[NSR];
  auto load_s = [&](int d) {
#pragma unroll
    for (int v = 0; v < NSR; ++v) sr[v] = sS[d * M2 + 16 * v + (tid & 15)];
    asm volatile("s_nop 1");
  };
  auto base_of = [&](auto dir_tag, int q) {
    constexpr int D = decltype(dir_tag)::value;
    return D == 0 ? P * pa[q] + PL * pb[q] : (D == 1 ? pa[q] + PL * pb[q] : pa[q] + P * pb[q]);
  };
  auto load_x = [&](auto dir_tag, double (&x)[PPT][NL]) {
    constexpr int D = decltype(dir_tag)::value;
    constexpr int st = D == 0 ? 1 : (D == 1 ? P : PL);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int b = base_of(dir_tag, q);
#pragma unroll
      for (int m = 0; m < NL; ++m) x[q][m] = u[b + st * m];
    }
  };
  auto store_y = [&](auto dir_tag, const double (&y)[PPT][NL]) {
    constexpr int D = decltype(dir_tag)::value;
    constexpr int st = D == 0 ? 1 : (D == 1 ? P : PL);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int b = base_of(dir_tag, q);
      if (pv[q]) {
#pragma unroll
        for (int m = 0; m < NL; ++m) u[b + st * m] = y[q][m];
      }
    }
  };
  using R_ = std::integral_constant<int, 0>;
  using S_ = std::integral_constant<int, 1>;
  using T_ = std::integral_constant<int, 2>;
  using Pairs = std::make_integer_sequence<int, (NL + 1) / 2>;
  // element order as in k_sch_fdm (face neighbours on one XCD together)
  const int64_t q8 = nel / 8, r8 = nel % 8;
  for (int64_t ep = blockIdx.x; ep < nel; ep += gridDim.x) {
    const int64_t x8 = ep % 8;
    const int64_t e = x8 * q8 + (x8 < r8 ? x8 : r8) + ep / 8;
    const double *re = r + e * N3;
    const int32_t *ge = gin + e * 6 * N2;
    const int32_t *go = gout + e * 6 * N2;
    const double *Se = S + e * 3 * M2;
    double vr[NR], vs[NS], vg[NG], vl;
    int32_t gi[NG], gd[NG];
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T, qc = q < 6 * N2 ? q : 6 * N2 - 1;
      gi[it] = q < 6 * N2 ? ge[qc] : -1;
      gd[it] = q < 6 * N2 ? go[qc] : -1;
    }
#pragma unroll
    for (int it = 0; it < NR; ++it) {
      const int q = tid + it * T;
      vr[it] = re[q < N3 ? q : N3 - 1];
    }
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * T;
      vs[it] = Se[q < 3 * M2 ? q : 3 * M2 - 1];
    }
    vl = lam[e * NLAM + (tid < NLAM ? tid : NLAM - 1)];
#pragma unroll
    for (int it = 0; it < NG; ++it) vg[it] = r[gi[it] >= 0 ? gi[it] : 0];
    __syncthreads();  // the previous element's LDS reads are done
#pragma unroll
    for (int it = 0; it < NR; ++it) {
      const int q = tid + it * T;
      if (q < N3) {
        const int i = q % N, j = (q / N) % N, k = q / N2;
        u[(i + 1) + P * (j + 1) + PL * (k + 1)] = vr[it];
      }
    }
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * T;
      if (q < 3 * M2) sS[q] = vs[it];
    }
    if (tid < 16) sS[3 * M2 + tid] = 0.0;
    if (tid < NLAM) sl[tid] = vl;
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T;
      if (q < 6 * N2) u[gpos(q)] = gi[it] >= 0 ? vg[it] : 0.0;
    }
    for (int q = tid; q < M2 * NL; q += T) {  // ghost edges and corners
      const int i = q % NL, j = (q / NL) % NL, k = q / M2;
      const int g = (i == 0 || i == NL - 1) + (j == 0 || j == NL - 1) + (k == 0 || k == NL - 1);
      if (g >= 2) u[i + P * j + PL * k] = 0.0;
    }
    __syncthreads();
    double x[PPT][NL], y[PPT][NL];
    load_s(0);
    load_x(R_(), x);
    fdm_matvec<NL, PPT, true>(y, sr, x, Pairs());
    store_y(R_(), y);
    __syncthreads();
    load_s(1);
    load_x(S_(), x);
    fdm_matvec<NL, PPT, true>(y, sr, x, Pairs());
    store_y(S_(), y);
    __syncthreads();
    load_s(2);
    load_x(T_(), x);
    fdm_matvec<NL, PPT, true>(y, sr, x, Pairs());
    // the diagonal (hsmg.f:753-770: 0 where lr + ls + lt <= eps)
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const double lij = sl[pa[q]] + sl[NL + pb[q]], eps = sl[3 * NL];
#pragma unroll
      for (int k = 0; k < NL; ++k) {
        const double d = lij + sl[2 * NL + k];
        double rc = __builtin_amdgcn_rcp(d);
        rc = fma(rc, fma(-d, rc, 1.0), rc);
        rc = fma(rc, fma(-d, rc, 1.0), rc);
        y[q][k] = (d > eps) ? y[q][k] * rc : 0.0;
      }
    }
    asm volatile("s_nop 1");
    fdm_matvec<NL, PPT, false>(x, sr, y, Pairs());
    store_y(T_(), x);
    __syncthreads();
    load_s(1);
    load_x(S_(), x);
    fdm_matvec<NL, PPT, false>(y, sr, x, Pairs());
    store_y(S_(), y);
    __syncthreads();
    load_s(0);
    load_x(R_(), x);
    fdm_matvec<NL, PPT, false>(y, sr, x, Pairs());
    store_y(R_(), y);
    __syncthreads();
    double *wi = wint + e * N3;
#pragma unroll
    for (int it = 0; it < NR; ++it) {
      const int q = tid + it * T;
      if (q < N3) {
        const int i = q % N, j = (q / N) % N, k = q / N2;
        wi[q] = u[(i + 1) + P * (j + 1) + PL * (k + 1)];
      }
    }
#pragma unroll
    for (int it = 0; it < NG; ++it) {
      const int q = tid + it * T;
      if (gd[it] >= 0) wgh[gd[it]] = u[gpos(q)];
    }
  }
}

// Overlap sum (hsmg.f:475-481) at node (i,j,k) of element el: the interior
// of the local solution, plus on the first interior layer the neighbour's
// solution at the ghost node that coincides with it (delivered to this
// element's ghost slots by k_sch_fdm; 0 where there is no neighbour), face by
// face in the order r-, r+, s-, s+, t-, t+.  k_sch_out, the fused
// prolongation (k_tnsr MODE 4) and the fused gather-scatter (k_gs_fin_ov)
// all form it here, in this order, so their values are identical.
template <int N, typename R>
__device__ __forceinline__ R overlap_value(const R *__restrict__ wint,
                                           const R *__restrict__ wgh, int64_t el, int i, int j,
                                           int k) {
  constexpr int N2 = N * N, N3 = N2 * N;
  const R *g = wgh + el * 6 * N2;
  R a = wint[el * N3 + i + N * (j + N * k)];
  if (i == 1) a += g[0 * N2 + j + N * k];
  if (i == N - 2) a += g[1 * N2 + j + N * k];
  if (j == 1) a += g[2 * N2 + i + N * k];
  if (j == N - 2) a += g[3 * N2 + i + N * k];
  if (k == 1) a += g[4 * N2 + i + N * j];
  if (k == N - 2) a += g[5 * N2 + i + N * j];
  return a;
}

// Overlap sum of every node into e (the path without the fused kernels),
// the sum of overlap_value in the same order.  One wavefront per element:
// the interior values and the six ghost faces are loaded coalesced and
// unconditionally (one round trip), the faces staged in LDS, where the
// first-interior-layer nodes pick up their neighbour values.
template <typename R, int N, typename EO = R>
__global__ void __launch_bounds__(64)
    k_sch_out(EO *__restrict__ eo, const R *__restrict__ wint, const R *__restrict__ wgh,
              int64_t nel) {
  constexpr int N2 = N * N, N3 = N2 * N, U = (N3 + 63) / 64, G = 6 * N2, UG = (G + 63) / 64;
  __shared__ R sg[G];
  const int l = (int)threadIdx.x;
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    R w[U], gv[UG];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = l + 64 * u;
      w[u] = q < N3 ? wint[e * N3 + q] : R(0);
    }
#pragma unroll
    for (int u = 0; u < UG; ++u) {
      const int q = l + 64 * u;
      gv[u] = q < G ? wgh[e * G + q] : R(0);
    }
    __syncthreads();  // the previous element's reads of sg are done
#pragma unroll
    for (int u = 0; u < UG; ++u) {
      const int q = l + 64 * u;
      if (q < G) sg[q] = gv[u];
    }
    __syncthreads();
    EO *out = eo + e * N3;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = l + 64 * u;
      if (q >= N3) continue;
      const int i = q % N, j = (q / N) % N, k = q / N2;
      R a = w[u];
      if (i == 1) a += sg[0 * N2 + j + N * k];
      if (i == N - 2) a += sg[1 * N2 + j + N * k];
      if (j == 1) a += sg[2 * N2 + i + N * k];
      if (j == N - 2) a += sg[3 * N2 + i + N * k];
      if (k == 1) a += sg[4 * N2 + i + N * j];
      if (k == N - 2) a += sg[5 * N2 + i + N * j];
      out[q] = EO(a);
    }
  }
}

template <typename R, typename EO = R>
int launch_sch_out(EO *e, const R *wint, const R *wgh, int nh, int64_t nel, hipStream_t s) {
  const int grid = (int)std::min<int64_t>(nel, 1 << 20);
  switch (nh) {
#define NK_O(N)                                                                             \
  case N:                                                                                   \
    hipLaunchKernelGGL((k_sch_out<R, N, EO>), dim3(grid), dim3(64), 0, s, e, wint, wgh, nel); \
    break;
    NK_O(3) NK_O(4) NK_O(5) NK_O(6) NK_O(7) NK_O(8) NK_O(9) NK_O(10) NK_O(11) NK_O(12)
    NK_O(13) NK_O(14) NK_O(15) NK_O(16)
#undef NK_O
    default:
      set_error("overlap sum of %d points not instantiated", nh);
      return NKH_ERR_UNSUPPORTED;
  }
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// Per-element tensor apply with one matrix in all directions,
// s = (A (x) A (x) A) u, A (NV x NU) column-major.
//   MODE 1 (restriction, hsmg.f:2216-2232): v = s with u weighted by w first;
//   MODE 2 (prolongation, hsmg.f:1930-1946, with the Schwarz weights of the
//           finer level): v = v * w + s.
//   MODE 3 (prolongation into the fine level before gs_fin): v = s on
//           element-boundary nodes, v = eb * w + s inside.
//   MODE 4: MODE 3 with eb formed from the local solutions (eb = wint, eg =
//           wgh, of type EB) by the overlap sum, rounded to EO as a stored
//           overlap sum would be.
template <int NV, int NU, int MODE, typename EB, typename UT, typename VT, typename WT,
          typename EO = EB>
__global__ void k_tnsr(VT *__restrict__ v, const UT *__restrict__ u,
                       const WT *__restrict__ w, const double *__restrict__ A, int64_t nel,
                       const EB *__restrict__ eb, const double *__restrict__ om,
                       const EB *__restrict__ eg) {
  constexpr int NV3 = NV * NV * NV, NU3 = NU * NU * NU;
  __shared__ double sA[NV * NU];
  __shared__ double su[NU3];
  __shared__ double s1[NV * NU * NU];
  __shared__ double s2[NV * NV * NU];
  for (int t = threadIdx.x; t < NV * NU; t += blockDim.x) sA[t] = A[t];
  for (int64_t e = blockIdx.x; e < nel; e += gridDim.x) {
    __syncthreads();
    for (int t = threadIdx.x; t < NU3; t += blockDim.x) {
      double x = double(u[(int64_t)e * NU3 + t]);
      if constexpr (MODE == 1) x = x * wdec(w, (int64_t)e * NU3 + t, nullptr);
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
      const int64_t o = (int64_t)e * NV3 + t;
      if constexpr (MODE == 2) {
        v[o] = VT(double(v[o]) * wdec(w, o, nullptr) + s);
      } else if constexpr (MODE == 3) {
        const int i = t % NV, j = (t / NV) % NV, k = t / (NV * NV);
        const bool bnd = (i == 0 || i == NV - 1 || j == 0 || j == NV - 1 || k == 0 || k == NV - 1);
        v[o] = VT(bnd ? s : double(eb[o]) * wdec(w, o, nullptr) + s);
      } else if constexpr (MODE == 4) {
        const int i = t % NV, j = (t / NV) % NV, k = t / (NV * NV);
        const bool bnd = (i == 0 || i == NV - 1 || j == 0 || j == NV - 1 || k == 0 || k == NV - 1);
        v[o] = VT(bnd ? s
                      : double(EO(overlap_value<NV, EB>(eb, eg, e, i, j, k))) *
                                wdec(w, o, nullptr) +
                            s);
      } else {
        v[o] = VT(om ? s * om[o] : s);
      }
    }
  }
}

// ---- coarse problem ------------------------------------------------------
// coarse right-hand side: b[slot] = sum of the level-1 values in that slot
// (crs_solve assembles by vertex id, fcrs.c:77-85)
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

__global__ void k_zero_int(int *flag) { *flag = 0; }

// the held coarse unknown's right-hand side: 0
__global__ void k_crs_pin(double *b, int pin) { b[pin] = 0.0; }

// x <- x - mean(x) over n entries, one block, fixed summation order
__global__ void __launch_bounds__(1024) k_sub_mean_1b(double *__restrict__ x, int n) {
  __shared__ double sw[16];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += x[i];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = a;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sw[w];
  const double m = s / (double)n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] -= m;
}

// ORTHO (navier1.f:223-255): partial sums of v over a fixed grid, then every
// block of the second kernel forms the same total and subtracts the mean
constexpr int kOrthoBlocks = 1024;
__global__ void __launch_bounds__(256) k_ortho_part(const double *__restrict__ v, int64_t n,
                                                    double *__restrict__ part) {
  __shared__ double sw[4];
  double a = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a += v[i];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
}
__global__ void __launch_bounds__(256) k_ortho_sub(double *__restrict__ v, int64_t n,
                                                   const double *__restrict__ part, int nb) {
  __shared__ double tot;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    tot = s / (double)n;
  }
  __syncthreads();
  const double m = tot;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] -= m;
}

// 1 where a != b (flag set, never cleared here)
__global__ void k_neq(const double *a, const double *b, int64_t n, int *flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (a[i] != b[i]) atomicOr(flag, 1);
}

// The local solves with the six contractions on the matrix cores: one wave
// per element; a contraction along one direction is the product of the
// element's NL x NL matrix (zero-padded to 16 x 16) with the NL x NL^2 array
// of pencils along that direction, in 16-column blocks.  Single precision
// (R = float): BF, bf16 operands (v_mfma_f32_16x16x16_bf16, one product per
// block), else FP32 operands (v_mfma_f32_16x16x4_f32, ceil(NL/4) products per
// block); FP32 accumulation either way.  Double precision (R = double): FP64
// operands and accumulation (v_mfma_f64_16x16x4_f64, ceil(NL/4) products per
// block).  Operand layouts (checked on gfx950): lane l holds A[l%16][k],
// B[k][l%16] with k = l/16 (FP32 and FP64, K = 4) or 4(l/16)+v (bf16,
// K = 16); register v of the result holds D[4(l/16)+v][l%16] for the FP32
// accumulators but D[4v+l/16][l%16] for the FP64 ones.  Assembly of the
// extended array, the diagonal and the write-back are those of k_sch_fdm.
// Kernel of the local solves (process-wide): 0 the vector kernels, 1 FP32
// matrix-core products for the single-precision smoother, 2 bf16 ones, 3
// FP64 matrix-core products for the double-precision smoother.
int g_fdm_variant = 0;

typedef short nk_v4s __attribute__((ext_vector_type(4)));
typedef float nk_v4f __attribute__((ext_vector_type(4)));
typedef double nk_v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short nk_bf16(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

template <int NL, typename R, typename RI, bool BF>
__global__ void __launch_bounds__(64)
    k_sch_fdm_mx(R *__restrict__ wint, R *__restrict__ wgh, const RI *__restrict__ r,
                 const int32_t *__restrict__ gin, const int32_t *__restrict__ gout,
                 const R *__restrict__ S, const R *__restrict__ lam, int64_t nel) {
  static_assert(NL <= 16, "one 16 x 16 tile per contraction");
  static_assert(!BF || std::is_same<R, float>::value, "bf16 operands accumulate in FP32");
  constexpr bool F64 = std::is_same<R, double>::value;
  using Acc = typename std::conditional<F64, nk_v4d, nk_v4f>::type;
  constexpr int N = NL - 2, N2 = N * N, N3 = N2 * N, M2 = NL * NL, NLAM = 3 * NL + 1;
  constexpr int P = NL | 1, PL = P * NL, T = 64, CB = (M2 + 15) / 16, KQ = (NL + 3) / 4;
  __shared__ R u[PL * NL];
  __shared__ R sS[3 * M2];
  __shared__ R sl[NLAM];
  const int tid = threadIdx.x, lr = tid & 15, lg = tid >> 4;
  auto gpos = [](int q) {
    const int f = q / N2, a = q % N, b = (q / N) % N;
    int i, j, k;
    switch (f) {
      case 0: i = 0; j = a + 1; k = b + 1; break;
      case 1: i = NL - 1; j = a + 1; k = b + 1; break;
      case 2: i = a + 1; j = 0; k = b + 1; break;
      case 3: i = a + 1; j = NL - 1; k = b + 1; break;
      case 4: i = a + 1; j = b + 1; k = 0; break;
      default: i = a + 1; j = b + 1; k = NL - 1; break;
    }
    return i + P * j + PL * k;
  };
  // y_i = sum_m M(i,m) x_m for every pencil along direction D, M(i,m) =
  // Sm(m + NL i) when TR else Sm(i + NL m)
  auto contract = [&](auto dir_tag, auto tr_tag, const R *Sm) {
    constexpr int D = decltype(dir_tag)::value;
    constexpr bool TR = decltype(tr_tag)::value;
    constexpr int st = D == 0 ? 1 : (D == 1 ? P : PL);
    auto Mv = [&](int i, int m) -> R {
      return (i < NL && m < NL) ? (TR ? Sm[m + NL * i] : Sm[i + NL * m]) : R(0);
    };
    nk_v4s ab;
    R af[KQ];
    if constexpr (BF) {
#pragma unroll
      for (int v = 0; v < 4; ++v) ab[v] = nk_bf16(Mv(lr, 4 * lg + v));
    } else {
#pragma unroll
      for (int q = 0; q < KQ; ++q) af[q] = Mv(lr, 4 * q + lg);
    }
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const int col = cb * 16 + lr;
      const bool cv = col < M2;
      const int pa = cv ? col % NL : 0, pb = cv ? col / NL : 0;
      const int base = D == 0 ? P * pa + PL * pb : (D == 1 ? pa + PL * pb : pa + P * pb);
      Acc acc = {R(0), R(0), R(0), R(0)};
      if constexpr (BF) {
        nk_v4s bb;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = 4 * lg + v;
          bb[v] = nk_bf16((cv && m < NL) ? u[base + st * m] : 0.0f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab, bb, acc, 0, 0, 0);
      } else {
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
          const int m = 4 * q + lg;
          const R bv = (cv && m < NL) ? u[base + st * m] : R(0);
          if constexpr (F64)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q], bv, acc, 0, 0, 0);
          else
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[q], bv, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = F64 ? 4 * v + lg : 4 * lg + v;
        if (cv && i < NL) u[base + st * i] = acc[v];
      }
    }
  };
  using R_ = std::integral_constant<int, 0>;
  using S_ = std::integral_constant<int, 1>;
  using T_ = std::integral_constant<int, 2>;
  using TR_ = std::true_type;
  using NT_ = std::false_type;
  const int64_t q8 = nel / 8, r8 = nel % 8;
  for (int64_t ep = blockIdx.x; ep < nel; ep += gridDim.x) {
    const int64_t x8 = ep % 8;
    const int64_t e = x8 * q8 + (x8 < r8 ? x8 : r8) + ep / 8;
    __syncthreads();
    const RI *re = r + e * N3;
    const int32_t *ge = gin + e * 6 * N2;
    const R *Se = S + e * 3 * M2;
    for (int q = tid; q < 3 * M2; q += T) sS[q] = Se[q];
    if (tid < NLAM) sl[tid] = lam[e * NLAM + tid];
    for (int q = tid; q < N3; q += T) {
      const int i = q % N, j = (q / N) % N, k = q / N2;
      u[(i + 1) + P * (j + 1) + PL * (k + 1)] = R(re[q]);
    }
    for (int q = tid; q < 6 * N2; q += T) {
      const int32_t sidx = ge[q];
      u[gpos(q)] = (sidx >= 0) ? R(r[sidx]) : R(0);
    }
    for (int q = tid; q < M2 * NL; q += T) {  // ghost edges and corners
      const int i = q % NL, j = (q / NL) % NL, k = q / M2;
      const int g = (i == 0 || i == NL - 1) + (j == 0 || j == NL - 1) + (k == 0 || k == NL - 1);
      if (g >= 2) u[i + P * j + PL * k] = R(0);
    }
    __syncthreads();
    contract(R_(), TR_(), sS);
    __syncthreads();
    contract(S_(), TR_(), sS + M2);
    __syncthreads();
    contract(T_(), TR_(), sS + 2 * M2);
    __syncthreads();
    for (int q = tid; q < M2 * NL; q += T) {
      const int i = q % NL, j = (q / NL) % NL, k = q / M2;
      const R lij = sl[i] + sl[NL + j], eps = sl[3 * NL];
      const R diag = lij + sl[2 * NL + k];
      const int a = i + P * j + PL * k;
      u[a] = (diag > eps) ? u[a] * (R(1) / diag) : R(0);
    }
    __syncthreads();
    contract(T_(), NT_(), sS + 2 * M2);
    __syncthreads();
    contract(S_(), NT_(), sS + M2);
    __syncthreads();
    contract(R_(), NT_(), sS);
    __syncthreads();
    R *wi = wint + e * N3;
    for (int q = tid; q < N3; q += T) {
      const int i = q % N, j = (q / N) % N, k = q / N2;
      wi[q] = u[(i + 1) + P * (j + 1) + PL * (k + 1)];
    }
    const int32_t *go = gout + e * 6 * N2;
    for (int q = tid; q < 6 * N2; q += T) {
      const int32_t d = go[q];
      if (d >= 0) wgh[d] = u[gpos(q)];
    }
  }
}

template <int NL, typename R, typename RI, bool BF>
void launch_sch_fdm_mx(R *wint, R *wgh, const R *S, const R *lam, const MgLevel &L,
                       const RI *r, int64_t nel, hipStream_t s) {
  const int grid = (int)std::min<int64_t>((nel + 7) / 8 * 8, 1 << 20);
  hipLaunchKernelGGL((k_sch_fdm_mx<NL, R, RI, BF>), dim3(grid), dim3(64), 0, s, wint, wgh, r,
                     L.gin, L.gout, S, lam, nel);
}

// NEKAMD_OVERLAP_FUSED=1 (read once) forms the fine level's overlap sum
// inside the prolongation and the final gather-scatter (k_tnsr MODE 4,
// k_gs_fin_ov) instead of in its own pass before the join.  Measured at
// E = 64^3, N = 7: 235.7 ms per FP64 solve fused, 232.0 separate (mixed
// 156.3 / 150.8): the fused kernels move work behind the coarse chain.
static bool overlap_fused() {
  static const bool v = [] {
    const char *e = std::getenv("NEKAMD_OVERLAP_FUSED");
    return e && std::atoi(e) != 0;
  }();
  return v;
}

template <int NL, typename R, typename RI>
void launch_sch_fdm(R *wint, R *wgh, const R *S, const R *lam, const MgLevel &L, const RI *r,
                    int64_t nel, hipStream_t s) {
  const int grid = (int)std::min<int64_t>((nel + 7) / 8 * 8, 1 << 20);
  if constexpr (std::is_same<R, double>::value && std::is_same<RI, double>::value)
    hipLaunchKernelGGL((k_sch_fdm_d<NL>), dim3(grid), dim3(Fdm<NL>::kThreads), 0, s, wint,
                       wgh, r, L.gin, L.gout, S, lam, nel);
  else
    hipLaunchKernelGGL((k_sch_fdm<NL, R, RI>), dim3(grid), dim3(Fdm<NL>::kThreads), 0, s,
                       wint, wgh, r, L.gin, L.gout, S, lam, nel);
}

}  // namespace

int mg_overlap_out(Mg &mg, int l, double *e, hipStream_t s) {
  MgLevel &L = mg.lev[l];
  NK_TRY(launch_sch_out<double>(e, L.wint, L.wgh, L.nh, mg.nel, s));
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// Local solves + overlap sum in precision R: double uses the level's own
// arrays, float the single-precision copies (fine level, mixed precision).
template <typename R, typename RI = double, typename EO = R>
static int schwarz_local(Mg &mg, int l, const RI *r, EO *e, hipStream_t s) {
  MgLevel &L = mg.lev[l];
  R *wint, *wgh;
  const R *S, *lam;
  if constexpr (std::is_same<R, double>::value) {
    wint = L.wint; wgh = L.wgh; S = L.fs; lam = L.fl;
  } else {
    NK_CHECK(L.fs32, NKH_ERR_STATE, "single-precision smoother data not prepared");
    wint = L.wint32; wgh = L.wgh32; S = L.fs32; lam = L.fl32;
  }
  bool done = false;
  const int mx = g_fdm_variant;
  constexpr bool f32 = std::is_same<R, float>::value;
  if ((f32 ? (mx == 1 || mx == 2) : mx == 3) && L.nh + 2 <= 16) {
    switch (L.nh + 2) {
#define NK_C(NL)                                                                \
  case NL:                                                                      \
    if constexpr (f32)                                                          \
      if (mx == 2) {                                                            \
        launch_sch_fdm_mx<NL, R, RI, true>(wint, wgh, S, lam, L, r, mg.nel, s);  \
        break;                                                                  \
      }                                                                         \
    launch_sch_fdm_mx<NL, R, RI, false>(wint, wgh, S, lam, L, r, mg.nel, s);     \
    break;
      NK_C(5) NK_C(6) NK_C(7) NK_C(8) NK_C(9) NK_C(10) NK_C(11) NK_C(12) NK_C(13) NK_C(14)
      NK_C(15) NK_C(16)
#undef NK_C
      default: break;
    }
    done = true;
  }
  if (!done) switch (L.nh + 2) {
#define NK_C(NL) \
  case NL: launch_sch_fdm<NL, R, RI>(wint, wgh, S, lam, L, r, mg.nel, s); break;
    NK_C(5) NK_C(6) NK_C(7) NK_C(8) NK_C(9) NK_C(10) NK_C(11) NK_C(12) NK_C(13) NK_C(14)
    NK_C(15) NK_C(16) NK_C(17) NK_C(18)
#undef NK_C
    default:
      set_error("Schwarz block of %d points not instantiated", L.nh + 2);
      return NKH_ERR_UNSUPPORTED;
  }
  NK_HIP(hipGetLastError());
  // e == nullptr: the overlap sum is left to the fused consumers
  if (e) NK_TRY((launch_sch_out<R, EO>(e, wint, wgh, L.nh, mg.nel, s)));
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// The last step of the V-cycle at the fine level (gs_fin's dssum + weights +
// coarse correction + dsavg, ic.f:1871-1897) with the Schwarz sum e of each
// member formed from the local solutions by the overlap sum (rounded to EO,
// the type a stored overlap sum would have) instead of read from memory.
// Same group walk and summation order as k_gs_fin.
template <int N, typename R, typename EO, typename Z, typename W>
__global__ void __launch_bounds__(256)
    k_gs_fin_ov(Z *__restrict__ z, const R *__restrict__ wint, const R *__restrict__ wgh,
                const W *__restrict__ w, const int32_t *__restrict__ off,
                const int32_t *__restrict__ idx, int64_t ngroups, int64_t chunk) {
  constexpr int N2 = N * N, N3 = N2 * N;
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t lb = (int64_t)(b % 8) * (nb / 8) + b / 8;
  const int64_t s0 = lb * chunk;
  const int64_t s1 = (s0 + chunk < ngroups) ? s0 + chunk : ngroups;
  auto ev = [&](int32_t l) {
    const int64_t el = l / N3;
    const int q = (int)(l - el * N3);
    return EO(overlap_value<N, R>(wint, wgh, el, q % N, (q / N) % N, q / N2));
  };
  for (int64_t g = s0 + threadIdx.x; g < s1; g += blockDim.x) {
    const int32_t a = off[g], c = off[g + 1];
    EO Se = ev(idx[a]);
    for (int32_t q = a + 1; q < c; ++q) Se += ev(idx[q]);
    const double S = double(Se);
    double acc = S * double(w[idx[a]]) + double(z[idx[a]]);
    for (int32_t q = a + 1; q < c; ++q) {
      const int32_t l = idx[q];
      acc += S * double(w[l]) + double(z[l]);
    }
    acc = acc * (1.0 / (double)(c - a));
    for (int32_t q = a; q < c; ++q) z[idx[q]] = Z(acc);
  }
}

template <typename R, typename EO, typename Z, typename W>
static int gs_fin_ov(const GsMap &gs, int nh, Z *z, const R *wint, const R *wgh, const W *w,
                     hipStream_t s) {
  if (gs.ngroups == 0) return NKH_OK;
  int64_t nb = (gs.ngroups + 511) / 512;
  nb = (nb + 7) / 8 * 8;
  nb = nb < 64 ? 64 : (nb > kGsBlocks ? kGsBlocks : nb);
  const int64_t chunk = (gs.ngroups + nb - 1) / nb;
  switch (nh) {
#define NK_G(N)                                                                               \
  case N:                                                                                     \
    hipLaunchKernelGGL((k_gs_fin_ov<N, R, EO, Z, W>), dim3((int)nb), dim3(256), 0, s, z, wint, \
                       wgh, w, gs.off, gs.idx, gs.ngroups, chunk);                            \
    break;
    NK_G(3) NK_G(4) NK_G(5) NK_G(6) NK_G(7) NK_G(8) NK_G(9) NK_G(10) NK_G(11) NK_G(12)
    NK_G(13) NK_G(14) NK_G(15) NK_G(16)
#undef NK_G
    default:
      set_error("fused overlap gather-scatter of %d points not instantiated", nh);
      return NKH_ERR_UNSUPPORTED;
  }
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int fdm_set_variant(int v) {
  NK_CHECK(v >= 0 && v <= 3, NKH_ERR_ARG, "FDM kernel variant %d (0, 1, 2 or 3)", v);
  g_fdm_variant = v;
  return NKH_OK;
}

int fdm_variant() { return g_fdm_variant; }

int mg_schwarz_local(Mg &mg, int l, const double *r, double *e, hipStream_t s) {
  return schwarz_local<double>(mg, l, r, e, s);
}

// om (MODE 1 only, may be null): 0/1 mask applied to the output.  UT/VT:
// storage of the input / output field (arithmetic in FP64).
template <int MODE, typename EB = double, typename UT = double, typename VT = double,
          typename WT = double, typename EO = EB>
static int run_tnsr(VT *v, int nv, const UT *u, int nu, const WT *w, const double *A,
                    int64_t nel, hipStream_t s, const EB *eb = nullptr,
                    const double *om = nullptr, const EB *eg = nullptr) {
  const int grid = (int)std::min<int64_t>(nel, 1 << 20);
#define NK_T(V, U)                                                                       \
  if (nv == V && nu == U) {                                                              \
    hipLaunchKernelGGL((k_tnsr<V, U, MODE, EB, UT, VT, WT, EO>), dim3(grid),            \
                       dim3(V * V * V > 256 ? 256 : 64),                                  \
                       0, s, v, u, w, A, nel, eb, om, eg);                               \
    NK_HIP(hipGetLastError());                                                           \
    return NKH_OK;                                                                       \
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

// Coarse solve (hsmg.f:1321-1354 -> crs_solve, fcrs.c:77-85): assemble by
// unknown, solve, scatter back (Dirichlet vertices 0).
// crs_mode 1 applies one AMG V-cycle instead (the reference's param(40) = 1
// coarse solver is likewise one pass of its AMG hierarchy); crs_mode 2, the
// V-cycles inside the mixed solver's FP32 inner solves, runs the coarse
// AMG-PCG only to kInnerCoarseRtol: flexible CG absorbs the inexact coarse
// correction, and the coarse chain is the critical path of the V-cycle.
// Measured at E = 64^3, N = 7 (mixed solve to 1e-8): 16 outer iterations
// for every tolerance from 1e-13 to 0.3; 1e-4: 5.1 coarse iterations per
// V-cycle and 163.8 ms per solve, 1e-3: 3.9 / 161.8, 1e-2: 2.7 / 159.4,
// 3e-2: 2.0 / 158.2, 0.1: 1.75 / 157.5, 0.3: 1.25 / 156.0.  (One AMG
// V-cycle without the PCG's step length, crs_mode 1, takes 20 iterations.)
constexpr double kInnerCoarseRtol = 0.1;
static int coarse_solve(Mg &mg, double *e1, const double *r1, int crs_mode, hipStream_t s) {
  const int n = mg.ncrs;
  if (n > 0) {
    hipLaunchKernelGGL(k_crs_gather, dim3(mg_grid(n)), dim3(256), 0, s, mg.cb, r1, mg.goff,
                       mg.gidx, n);
    NK_HIP(hipGetLastError());
    if (mg.null_space) hipLaunchKernelGGL(k_crs_pin, dim3(1), dim3(1), 0, s, mg.cb, n - 1);
    if (crs_mode == 1) NK_TRY(mg.crs.vcycle_only(mg.cx, mg.cb, s));
    else if (crs_mode == 2) NK_TRY(mg.crs.solve(mg.cx, mg.cb, s, kInnerCoarseRtol));
    else NK_TRY(mg.crs.solve(mg.cx, mg.cb, s));
    if (mg.null_space)
      hipLaunchKernelGGL(k_sub_mean_1b, dim3(1), dim3(1024), 0, s, mg.cx, n);
  }
  hipLaunchKernelGGL(k_crs_scatter, dim3(mg_grid(8 * mg.nel)), dim3(256), 0, s, e1, mg.cx,
                     mg.slot, 8 * mg.nel);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// hsmg.f:1855-1949 H1MG_SOLVE (additive V-cycle, sigma = 1): z = M r, r
// already masked with the fine-level mask.
// The V-cycle with the fine-level input r and output z stored as FT: double
// (the drop-in; precision 1 runs the fine local solves in FP32 inside), or
// float (the single-precision inner solve of the mixed solver: fine level
// in FP32 throughout; the local solves of levels 2.. in FP32 between FP64
// residuals and corrections; the coarse solve in FP64).
template <typename FT>
static int mg_solve_t(Mesh &mesh, FT *z, const FT *r, hipStream_t s, int crs_mode,
                      int precision) {
  constexpr bool f32io = std::is_same<FT, float>::value;
  NK_CHECK(mesh.mg, NKH_ERR_STATE, "mesh has no multigrid hierarchy (nkh_mg_setup)");
  Mg &mg = *mesh.mg;
  const int Lm = mg.lmax;
  const int64_t nel = mg.nel;
  hipStream_t b = mg.sb;
  // fine level on s: local solves and overlap sum.  With ids zero exactly
  // at element interiors, the dssum of the Schwarz sum, the weights, the
  // coarse correction and dsavg are done together after the join (gs_fin);
  // otherwise dssum now and the plain sequence after the join.
  MgLevel &T = mg.lev[Lm];
  const bool merged = mesh.gs.interior_zero;
  const bool fuse_ov = merged && overlap_fused();
  const bool timed = timers().on;
  if (timed) NK_HIP(hipEventRecord(mg.tev[2], s));
  NK_CHECK(merged || !f32io, NKH_ERR_UNSUPPORTED,
           "single-precision V-cycle needs global ids zero exactly at element interiors");
  const bool fp32 = (precision == 1 || f32io) && merged;
  if (fp32) NK_TRY(mg_prepare_fp32(mg, s, f32io));
  if (fp32 && !fuse_ov && !T.e32)  // single-precision overlap sum, formed on its own
    NK_HIP(hipMalloc(&T.e32, sizeof(float) * std::max<int64_t>(T.n, 1)));
  // The streaming head of the coarse chain runs first, on s: the
  // restriction of the fine residual, its dssum and (when that level is 2)
  // the restriction to level 1.  Beside the fine local solves these
  // bandwidth-bound passes are starved of the compute units the solves
  // occupy, and the whole coarse chain waits on them; once they are done
  // the coarse solve, latency-bound, overlaps the fine level.
  MgLevel &L1 = mg.lev[1], &L2 = mg.lev[2];
  const double *rsrc = nullptr;  // restriction source below the fine level
  bool l1_restricted = false;
  if (Lm - 1 >= 2) {
    MgLevel &C = mg.lev[Lm - 1];
    // mask * dssum(v) = dssum(mask * v): the level mask is 0/1 and equal at
    // every copy of a node (made so by a product gs at set-up)
    if constexpr (f32io)  // single-precision input: single-precision weights
      NK_TRY((run_tnsr<1, double, FT, double, float>(C.r, C.nh, r, T.nh, T.rwt32, C.Jt, nel,
                                                     s, (const double *)nullptr, C.mask)));
    else if (T.rwt8)  // the weights 1/count as count bytes
      NK_TRY((run_tnsr<1, double, FT, double, CountCode>(
          C.r, C.nh, r, T.nh, reinterpret_cast<const CountCode *>(T.rwt8), C.Jt, nel, s,
          (const double *)nullptr, C.mask)));
    else
      NK_TRY((run_tnsr<1, double, FT, double>(C.r, C.nh, r, T.nh, T.rwt, C.Jt, nel, s,
                                              (const double *)nullptr, C.mask)));
    NK_TRY(gs_op(C.gs, C.r, 0, s));
    NK_HIP(hipEventRecord(mg.ev_lr[Lm - 1], s));
    rsrc = C.r;
    if (Lm - 1 == 2) {
      NK_TRY(run_tnsr<1>(L1.r, L1.nh, C.r, L2.nh, L2.rwt, L1.Jt, nel, s));
      l1_restricted = true;
    }
  }
  NK_HIP(hipEventRecord(mg.ev_fork, s));
  NK_HIP(hipStreamWaitEvent(b, mg.ev_fork, 0));
  // the local solves and their overlap sum (into T.e / T.e32; with
  // fuse_ov the sum is formed after the join from the local solutions)
  // (the timer events bracket the local solves alone)
  auto fine = [&]() -> int {
    if (timed) NK_HIP(hipEventRecord(mg.tev[0], s));
    if (fp32) {
      NK_TRY((schwarz_local<float, FT>(mg, Lm, r, (float *)nullptr, s)));
      if (timed) NK_HIP(hipEventRecord(mg.tev[1], s));
      if (!fuse_ov)
        NK_TRY((launch_sch_out<float>(T.e32, T.wint32, T.wgh32, T.nh, nel, s)));
    } else if constexpr (!f32io) {
      NK_TRY((schwarz_local<double, FT>(mg, Lm, r, (double *)nullptr, s)));
      if (timed) NK_HIP(hipEventRecord(mg.tev[1], s));
      if (!fuse_ov)
        NK_TRY((launch_sch_out<double>(merged ? T.e : z, T.wint, T.wgh, T.nh, nel, s)));
      if (!merged) NK_TRY(gs_op(T.gs, z, 0, s));
    }
    return NKH_OK;
  };
  // the Schwarz solve of intermediate level l on c (it only reads the
  // masked residual of its level)
  hipStream_t c = mg.sc;
  auto level = [&](int l, hipEvent_t after) -> int {
    MgLevel &C = mg.lev[l];
    NK_HIP(hipStreamWaitEvent(c, after, 0));
    if constexpr (f32io)  // inner solve of the mixed solver: FP32 local solves
      NK_TRY((schwarz_local<float, double, double>(mg, l, C.r, C.e, c)));
    else
      NK_TRY(mg_schwarz_local(mg, l, C.r, C.e, c));
    NK_TRY(gs_op(C.gs, C.e, 0, c));
    NK_HIP(hipEventRecord(mg.ev_le[l], c));
    return NKH_OK;
  };
  NK_TRY(fine());
  // coarser levels on b: restrict, dssum, mask; then the coarse solve
  for (int l = Lm - 1; l >= 2; --l) {
    MgLevel &C = mg.lev[l], &F = mg.lev[l + 1];
    if (l < Lm - 1) {  // (the level below the fine one was done on s above)
      NK_TRY(run_tnsr<1>(C.r, C.nh, rsrc, F.nh, F.rwt, C.Jt, nel, b, (const double *)nullptr,
                         C.mask));
      NK_TRY(gs_op(C.gs, C.r, 0, b));
      NK_HIP(hipEventRecord(mg.ev_lr[l], b));
    }
    NK_TRY(level(l, mg.ev_lr[l]));
    rsrc = C.r;
  }
  // the level-1 mask needs no pass of its own: the coarse gather reads only
  // unmasked vertices and the scatter writes 0 at masked ones (their slot is
  // -1), which is what masking r1 before and e1 after the solve would give
  if (Lm == 2)
    NK_TRY((run_tnsr<1, double, FT, double>(L1.r, L1.nh, r, L2.nh, L2.rwt, L1.Jt, nel, b)));
  else if (!l1_restricted)
    NK_TRY(run_tnsr<1>(L1.r, L1.nh, rsrc, L2.nh, L2.rwt, L1.Jt, nel, b));
  NK_TRY(coarse_solve(mg, L1.e, L1.r, crs_mode, b));
  for (int l = 2; l < Lm; ++l) {  // e_l = e_l * (mask*wt) + J e_{l-1}
    MgLevel &C = mg.lev[l - 1], &F = mg.lev[l];
    NK_HIP(hipStreamWaitEvent(b, mg.ev_le[l], 0));
    NK_TRY(run_tnsr<2>(F.e, F.nh, C.e, C.nh, F.mswt, C.J, nel, b));
  }
  NK_HIP(hipEventRecord(mg.ev_join, b));
  NK_HIP(hipStreamWaitEvent(s, mg.ev_join, 0));
  // z = dssum(e)*(mask*wt) + J e_{lmax-1}, then dsavg (ic.f:1871-1897)
  MgLevel &C = mg.lev[Lm - 1];
  if (fp32 && fuse_ov) {
    if constexpr (f32io) {
      NK_TRY((run_tnsr<4, float, double, FT, float>(z, T.nh, C.e, C.nh, T.mswt32, C.J, nel, s,
                                                    T.wint32, nullptr, T.wgh32)));
      NK_TRY((gs_fin_ov<float, float>(mesh.gs, T.nh, z, T.wint32, T.wgh32, T.mswt32, s)));
    } else {
      NK_TRY((run_tnsr<4, float, double, FT>(z, T.nh, C.e, C.nh, T.mswt, C.J, nel, s, T.wint32,
                                             nullptr, T.wgh32)));
      NK_TRY((gs_fin_ov<float, float>(mesh.gs, T.nh, z, T.wint32, T.wgh32, T.mswt, s)));
    }
  } else if (fp32) {
    if constexpr (f32io) {
      NK_TRY((run_tnsr<3, float, double, FT, float>(z, T.nh, C.e, C.nh, T.mswt32, C.J, nel, s,
                                                    T.e32)));
      // byte codes into the FP32-rounded half of the table: the same weights
      if (T.mswt8)
        NK_TRY(gs_fin(mesh.gs, z, T.e32, T.mswt8, mg.wlut + 256, s));
      else
        NK_TRY(gs_fin(mesh.gs, z, T.e32, T.mswt32, s));
    } else {
      NK_TRY((run_tnsr<3, float, double, FT>(z, T.nh, C.e, C.nh, T.mswt, C.J, nel, s, T.e32)));
      NK_TRY(gs_fin(mesh.gs, z, T.e32, T.mswt, s));
    }
  } else if constexpr (!f32io) {
    if (fuse_ov) {
      NK_TRY(run_tnsr<4>(z, T.nh, C.e, C.nh, T.mswt, C.J, nel, s, T.wint, nullptr, T.wgh));
      NK_TRY((gs_fin_ov<double, double>(mesh.gs, T.nh, z, T.wint, T.wgh, T.mswt, s)));
    } else if (merged) {
      NK_TRY(run_tnsr<3>(z, T.nh, C.e, C.nh, T.mswt, C.J, nel, s, T.e));
      // the final sum reads the weights as byte codes where every one has a
      // code (1 B instead of 8 per node).  In the prolongation above the
      // codes saved nothing (571 us with the weight decoded by arithmetic
      // from the code, slower through the table), so it keeps FP64 weights
      if (T.mswt8)
        NK_TRY(gs_fin(mesh.gs, z, T.e, T.mswt8, mg.wlut, s));
      else
        NK_TRY(gs_fin(mesh.gs, z, T.e, T.mswt, s));
    } else {
      NK_TRY(run_tnsr<2>(z, T.nh, C.e, C.nh, T.mswt, C.J, nel, s));
      NK_TRY(gs_op(mesh.gs, z, 4, s));
    }
  }
  if (timed) NK_HIP(hipEventRecord(mg.tev[3], s));
  return NKH_OK;
}

int mg_ortho(Mg &mg, double *v, int64_t n, hipStream_t s) {
  if (!mg.opart) NK_HIP(hipMalloc(&mg.opart, sizeof(double) * kOrthoBlocks));
  hipLaunchKernelGGL(k_ortho_part, dim3(kOrthoBlocks), dim3(256), 0, s, v, n, mg.opart);
  hipLaunchKernelGGL(k_ortho_sub, dim3(kOrthoBlocks), dim3(256), 0, s, v, n, mg.opart,
                     kOrthoBlocks);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int mg_solve(Mesh &mesh, double *z, const double *r, hipStream_t s, int crs_mode,
             int precision) {
  return mg_solve_t(mesh, z, r, s, crs_mode, precision);
}

int mg_solve32(Mesh &mesh, float *z, const float *r, hipStream_t s, int crs_mode) {
  return mg_solve_t(mesh, z, r, s, crs_mode, 1);
}

int mg_mask_fine(Mesh &mesh, double *r, hipStream_t s) {
  NK_CHECK(mesh.mg, NKH_ERR_STATE, "mesh has no multigrid hierarchy (nkh_mg_setup)");
  MgLevel &T = mesh.mg->lev[mesh.mg->lmax];
  return mul_inplace(r, T.mask, T.n, s);
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

// r = res; res = 0; p = 0; partial glsc3(r, wt, r).  r is stored masked
// with the multigrid's fine mask: the first h1mg_solve masks it in place
// before any other use (hsmg.f:1878 via h1mg_schwarz).
__global__ void __launch_bounds__(256)
    k_fcg_init(double *__restrict__ r, double *__restrict__ res, double *__restrict__ p,
               const double *__restrict__ wt, const double *__restrict__ mgmask, int64_t n,
               double *part) {
  double a = 0.0;
  NK_GRID_LOOP(i, n) {
    const double ri = res[i];
    r[i] = ri * mgmask[i];
    res[i] = 0.0;
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
    st->alpha = 0.0;  // the first deferred x update adds 0 * p (p = 0)
    st->iter = 0;
    st->done = 0;
    st->rnorm = div0;
  }
}

// The multiplicity weights wt (Nek's vmult = 1/dssum(1)) read from a byte
// per node: code c holds wt = 1/c exactly (1/c in FP64 is the value Nek
// computes), 1/8 of the weight array's traffic in the two streaming kernels
// of every iteration.  k_wt_code builds the codes and flags any weight that
// is not 1/c for a count c in 1..255 (then the FP64 weights are read).
__global__ void k_wt_code(const double *__restrict__ wt, uint8_t *__restrict__ code, int64_t n,
                          int *bad) {
  int b = 0;
  NK_GRID_LOOP(i, n) {
    const double w = wt[i];
    const double c = w > 0.0 ? rint(1.0 / w) : 0.0;
    const bool ok = c >= 1.0 && c <= 255.0 && 1.0 / c == w;
    code[i] = ok ? (uint8_t)c : (uint8_t)0;
    b |= ok ? 0 : 1;
  }
  if (b) atomicOr(bad, 1);
}
__device__ __forceinline__ double wt_at(const double *__restrict__ w, int64_t i) {
  return __builtin_nontemporal_load(w + i);
}
__device__ __forceinline__ double wt_at(const uint8_t *__restrict__ c, int64_t i) {
  return 1.0 / (double)c[i];
}

// partials of glsc3(z,wt,r) and glsc3(z,wt,w).  The reference keeps the
// previous residual r1 and forms rho2 = -glsc3(z,wt,r1-r) (hmholtz.f:2236-
// 2241); r1 - r is the last update alpha*w exactly in exact arithmetic (w
// still holds that iteration's A p), so rho2 = -alpha*glsc3(z,wt,w) without
// the extra vector.
template <typename WT>
__global__ void __launch_bounds__(256)
    k_fcg_dots(const double *__restrict__ z, const double *__restrict__ r,
               const double *__restrict__ w, const WT *__restrict__ wt, int64_t n,
               double *pa, double *pb) {
  double a = 0.0, b = 0.0;
  NK_GRID_LOOP(i, n) {
    const double zw = z[i] * wt_at(wt, i);
    a += zw * r[i];
    b += zw * w[i];
  }
  block_part2(a, b, pa, pb);
}

__global__ void __launch_bounds__(256) k_fcg_beta(FcgState *st, const double *pa, const double *pb) {
  const double a = final_sum(pa);
  const double b = final_sum(pb);
  if (threadIdx.x == 0) {
    st->rho0 = st->rho1;
    st->rho1 = a;
    // first iteration: r1 = 0, so rho2 = glsc3(z,wt,r); p = 0 then anyway
    st->rho2 = (st->iter == 0) ? a : -st->alpha * b;
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

// den from the count rho partials of the fused axhelm; alpha
// one block of 1024 threads (the fused axhelm leaves up to 65536 partials)
__global__ void __launch_bounds__(1024) k_fcg_alpha_n(FcgState *st, const double *part, int count) {
  const double t = block_sum_partials(part, count);
  if (threadIdx.x == 0) {
    st->den = t;
    st->alpha = st->rho1 / t;
  }
}

__global__ void __launch_bounds__(256) k_fcg_alpha(FcgState *st, const double *part) {
  const double t = final_sum(part);
  if (threadIdx.x == 0) {
    st->den = t;
    st->alpha = st->rho1 / t;
  }
}

// res += alpha p; r -= alpha w; partial sum r*r*wt
template <typename WT>
__global__ void __launch_bounds__(256)
    k_fcg_update(double *__restrict__ res, double *__restrict__ r, const double *__restrict__ p,
                 const double *__restrict__ w, const WT *__restrict__ wt, const FcgState *st,
                 int64_t n, double *part) {
  const double alpha = st->alpha;
  double a = 0.0;
  // streaming: every word is touched once, hence the nontemporal hints
  NK_GRID_LOOP(i, n) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(res + i) +
                                    alpha * __builtin_nontemporal_load(p + i),
                                res + i);
    const double ri =
        __builtin_nontemporal_load(r + i) - alpha * __builtin_nontemporal_load(w + i);
    __builtin_nontemporal_store(ri, r + i);
    a = a + ri * ri * wt_at(wt, i);
  }
  block_part(a, part);
}

// r -= alpha w; partial sum r*r*wt (the fused path: x += alpha p is done by
// the next iteration's axhelm, or by k_fcg_xfin after the last one)
template <typename WT>
__global__ void __launch_bounds__(256)
    k_fcg_update_r(double *__restrict__ r, const double *__restrict__ w,
                   const WT *__restrict__ wt, const FcgState *st, int64_t n, double *part) {
  const double alpha = st->alpha;
  double a = 0.0;
  NK_GRID_LOOP(i, n) {
    const double ri =
        __builtin_nontemporal_load(r + i) - alpha * __builtin_nontemporal_load(w + i);
    __builtin_nontemporal_store(ri, r + i);
    a = a + ri * ri * wt_at(wt, i);
  }
  block_part(a, part);
}

// the last iteration's x += alpha p
__global__ void k_fcg_xfin(double *__restrict__ x, const double *__restrict__ p,
                           const FcgState *st, int64_t n) {
  const double alpha = st->alpha;
  NK_GRID_LOOP(i, n) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) +
                                    alpha * __builtin_nontemporal_load(p + i),
                                x + i);
  }
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

int mg_wt_codes(Mg &mg, const double *wt, int64_t n, hipStream_t s, const uint8_t **codes) {
  // the flag after the n code bytes
  const int64_t wt8_flag = (n + 7) / 8 * 8;
  if (!mg.wt8) NK_HIP(hipMalloc(&mg.wt8, (size_t)wt8_flag + sizeof(int)));
  int *wt8_bad = reinterpret_cast<int *>(mg.wt8 + wt8_flag);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, wt8_bad);
  hipLaunchKernelGGL(k_wt_code, dim3(mg_grid(n)), dim3(256), 0, s, wt, mg.wt8, n, wt8_bad);
  NK_HIP(hipGetLastError());
  int wbad = 1;
  NK_HIP(hipMemcpyAsync(&wbad, wt8_bad, sizeof(int), hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  *codes = wbad ? nullptr : mg.wt8;
  return NKH_OK;
}

int run_flex_cg(Mesh &m, double *res, Coef h1, Coef h2, const double *wt, const double *mask,
                int maxit, const nkh_params &prm, nkh_result *out, double *rnorm_history) {
  NK_CHECK(m.mg, NKH_ERR_STATE, "hmh_flex_cg needs the multigrid hierarchy (nkh_mg_setup)");
  NK_CHECK(prm.param40 == 0 || prm.param40 == 1, NKH_ERR_UNSUPPORTED,
           "hmh_flex_cg: coarse solver param40=%d (0 exact, 1 one AMG V-cycle)", prm.param40);
  Mg &mg = *m.mg;
  hipStream_t s = cur_stream();
  const int64_t n = m.n;
  // (the mixed solver allocates fr..fw on its own, so each piece is checked)
  if (!mg.fr)
    for (double **p : {&mg.fr, &mg.fp, &mg.fz, &mg.fw}) NK_TRY(mg_dalloc(p, n));
  if (!mg.fpart) NK_TRY(mg_dalloc(&mg.fpart, 3 * kFcgBlocks));
  if (!mg.fst) NK_HIP(hipMalloc(&mg.fst, sizeof(FcgState)));
  if (!mg.fflag) NK_HIP(hipMalloc(&mg.fflag, sizeof(int)));
  const double *mgmask = mg.lev[mg.lmax].mask;
  double *r = mg.fr, *p = mg.fp, *z = mg.fz, *w = mg.fw;
  double *pa = mg.fpart, *pb = mg.fpart + kFcgBlocks;
  // the operator's kernel choice for this order (w is scratch until the
  // first apply of the iteration)
  NK_TRY(axhelm_autotune(m, w, res, h1, h2, s));
  // hmholtz.f:2205-2208: tolps from chktcg1 (pmask, vmult), capped by param(21)
  double tolps = 0.0;
  NK_TRY(run_chktcg1(m, &tolps, res, h1, h2, mask, wt));
  if (prm.param21 > 0 && tolps > std::fabs(prm.param21)) tolps = std::fabs(prm.param21);
  if (prm.istep == 0) tolps = 1.e-4;
  const dim3 g(kFcgBlocks), b(256);
  // r stays masked by the update (w carries pmask) when pmask is the
  // multigrid's fine mask; otherwise it is masked again before every V-cycle
  // as h1mg_solve does
  int differ = 0;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, mg.fflag);
  hipLaunchKernelGGL(k_neq, dim3(mg_grid(n)), b, 0, s, mask, mgmask, n, mg.fflag);
  NK_HIP(hipMemcpyAsync(&differ, mg.fflag, sizeof(int), hipMemcpyDeviceToHost, s));
  // the weights as counts, when every one is 1/count
  const uint8_t *wtc = nullptr;
  NK_TRY(mg_wt_codes(mg, wt, n, s, &wtc));
  // den inside the fused axhelm (p . A_L p) and the mask inside the
  // assembly, when the conditions for that identity hold (p = beta p + z
  // is continuous and zero where pmask is when pmask is the fine mask)
  bool fused = false;
  if (!differ) NK_TRY(fused_conditions(m, mask, wt, nullptr, &fused, s));
  double *rho_part = m.red + kRedRho1;
  const int nrho = axhelm_cg_grid(m);
  hipLaunchKernelGGL(k_fcg_init, g, b, 0, s, r, res, p, wt, mgmask, n, pa);
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
    if (differ && (rc = mul_inplace(r, mgmask, n, s)) != NKH_OK) break;
    if ((rc = mg_solve(m, z, r, s, prm.param40, prm.precision)) != NKH_OK) break;
    if (wtc) hipLaunchKernelGGL(k_fcg_dots<uint8_t>, g, b, 0, s, z, r, w, wtc, n, pa, pb);
    else hipLaunchKernelGGL(k_fcg_dots<double>, g, b, 0, s, z, r, w, wt, n, pa, pb);
    hipLaunchKernelGGL(k_fcg_beta, dim3(1), b, 0, s, mg.fst, pa, pb);
    if (fused) {
      if (tm.on) (void)hipEventRecord(ev[0], s);
      if ((rc = launch_axhelm_flex(m, w, p, z, h1, h2, &mg.fst->beta, rho_part, s, res,
                                   &mg.fst->alpha)) != NKH_OK)
        break;
      if (tm.on) (void)hipEventRecord(ev[1], s);
      if ((rc = gs_assemble(m.gs, w, nullptr, s)) != NKH_OK) break;
      if (tm.on) (void)hipEventRecord(ev[2], s);
      hipLaunchKernelGGL(k_fcg_alpha_n, dim3(1), dim3(1024), 0, s, mg.fst, rho_part, nrho);
    } else {
      hipLaunchKernelGGL(k_fcg_p, dim3(mg_grid(n)), b, 0, s, p, z, mg.fst, n);
      if (tm.on) (void)hipEventRecord(ev[0], s);
      if ((rc = launch_axhelm(m, w, p, h1, h2, s)) != NKH_OK) break;
      if (tm.on) (void)hipEventRecord(ev[1], s);
      if ((rc = gs_op(m.gs, w, 0, s)) != NKH_OK) break;
      if (tm.on) (void)hipEventRecord(ev[2], s);
      hipLaunchKernelGGL(k_fcg_den, g, b, 0, s, w, mask, wt, p, n, pa);
      hipLaunchKernelGGL(k_fcg_alpha, dim3(1), b, 0, s, mg.fst, pa);
    }
    if (fused && wtc)
      hipLaunchKernelGGL(k_fcg_update_r<uint8_t>, g, b, 0, s, r, w, wtc, mg.fst, n, pa);
    else if (fused)
      hipLaunchKernelGGL(k_fcg_update_r<double>, g, b, 0, s, r, w, wt, mg.fst, n, pa);
    else if (wtc)
      hipLaunchKernelGGL(k_fcg_update<uint8_t>, g, b, 0, s, res, r, p, w, wtc, mg.fst, n, pa);
    else
      hipLaunchKernelGGL(k_fcg_update<double>, g, b, 0, s, res, r, p, w, wt, mg.fst, n, pa);
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
      if (m.mg) {
        (void)hipEventElapsedTime(&a, m.mg->tev[0], m.mg->tev[1]);
        (void)hipEventElapsedTime(&c, m.mg->tev[2], m.mg->tev[3]);
        tm.tfdm += a * 1e-3;
        tm.nfdm += 1;
        tm.tmgv += c * 1e-3;
        tm.nmgv += 1;
      }
    }
    iter = hs.iter;
    if (rnorm_history) rnorm_history[k - 1] = hs.rnorm;
    if (hs.done) break;
  }
  if (tm.on)
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  NK_TRY(rc);
  // the fused path defers each x += alpha p to the next iteration's axhelm:
  // the last one is applied here
  if (fused && iter > 0) {
    hipLaunchKernelGGL(k_fcg_xfin, g, b, 0, s, res, p, mg.fst, n);
    NK_HIP(hipGetLastError());
  }
  // hmholtz.f:2272: orthogonalise the solution against the null space
  if (mg_null_space_on(mg, h2)) NK_TRY(mg_ortho(mg, res, n, s));
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
