// Jacobi-preconditioned CG (hmholtz.f:611-846 CGGO) and its helpers
// (hmholtz.f:527-609 CHKTCG1, :380-524 SETPREC), all data resident on the GPU.
//
// One CG iteration of the reference,
//     z = r*d ; rtz1 = (z,r)_mult ; rbn2 = sqrt((r,r)_{binv*mult}/vol)
//     exit test ; beta ; p = z + beta p ; w = A p ; dssum(w) ; w *= mask
//     rho = (w,p)_mult ; alpha ; x += alpha p ; r -= alpha w
// runs as five launches:
//     k_cg_update_dots : x/r update of the previous iteration + both dots
//     k_cg_scalars1    : one block: sums the dot partials, exit test, beta
//     axhelm (MODE 1)  : p update fused in, w = A_L p, rho = sum p.w
//     gs_assemble      : w <- mask * dssum(w)
//     k_cg_scalars2    : one block: rho, alpha
// rho is taken before the assembly: with p continuous and zero where the
// mask is, and mult summing to one over each group of shared nodes,
// (mask*QQ^T w, p)_mult = sum_l p_l w_l exactly in exact arithmetic.  The
// solve checks those conditions once and otherwise runs the unfused
// sequence (axhelm, dssum, mask, dot).
// Scalars stay on the device (CgState); the host enqueues iterations ahead
// and reads the state back every few iterations.  Kernels of iterations past
// the exit return immediately.
#include <cmath>
#include <cstring>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

namespace {

constexpr int kMaxCg = kMaxCgIter;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum of `count` partials by one block of kScalThreads threads in a fixed
// order.  Result valid on thread 0.
constexpr int kScalThreads = 1024;
__device__ double block_sum_array(const double *__restrict__ part, int count) {
  return block_sum_partials(part, count);
}

__device__ __forceinline__ void block_store2(double a, double b, double *pa, double *pb) {
  __shared__ double sa[4], sbb[4];
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sbb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = ((sa[0] + sa[1]) + sa[2]) + sa[3];
    pb[blockIdx.x] = ((sbb[0] + sbb[1]) + sbb[2]) + sbb[3];
  }
}

// Grid of the update-and-dots kernels: about 2048 nodes per block, at most
// kUpdBlocks -- a launch costs about 10 ns per block, which dominates small
// fields at a fixed large grid.  One partial sum per block.
inline int upd_blocks(int64_t n) {
  int64_t b = (n + 2047) / 2048;
  return (int)(b < 64 ? 64 : (b > kUpdBlocks ? kUpdBlocks : b));
}

// ---- generic reductions (setup) ----------------------------------------
// kind 0: sum a*b*c (c may be null -> 1, b may be null -> 1)
// kind 1: max |a|   kind 2: max a   kind 3: min a
__global__ void __launch_bounds__(256)
    k_reduce(const double *__restrict__ a, const double *__restrict__ b,
             const double *__restrict__ c, int64_t n, int kind, double *__restrict__ part) {
  __shared__ double sb[4];
  double v = (kind == 0 || kind == 1) ? 0.0 : (kind == 2 ? -INFINITY : INFINITY);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = a[i];
    if (kind == 0) {
      double t = x;
      if (b) t = t * b[i];
      if (c) t = t * c[i];
      v += t;
    } else if (kind == 1) {
      v = fmax(v, fabs(x));
    } else if (kind == 2) {
      v = fmax(v, x);
    } else {
      v = fmin(v, x);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_down(v, o, 64);
    if (kind == 0) v += t;
    else if (kind == 3) v = fmin(v, t);
    else v = fmax(v, t);
  }
  if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sb[0];
    for (int q = 1; q < 4; ++q) {
      if (kind == 0) t += sb[q];
      else if (kind == 3) t = fmin(t, sb[q]);
      else t = fmax(t, sb[q]);
    }
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(256)
    k_reduce_final(const double *__restrict__ part, int count, int kind,
                   double *__restrict__ out) {
  __shared__ double sb[4];
  double v = (kind == 0 || kind == 1) ? 0.0 : (kind == 2 ? -INFINITY : INFINITY);
  for (int q = threadIdx.x; q < count; q += blockDim.x) {
    const double x = part[q];
    if (kind == 0) v += x;
    else if (kind == 3) v = fmin(v, x);
    else v = fmax(v, x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_down(v, o, 64);
    if (kind == 0) v += t;
    else if (kind == 3) v = fmin(v, t);
    else v = fmax(v, t);
  }
  if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sb[0];
    for (int q = 1; q < 4; ++q) {
      if (kind == 0) t += sb[q];
      else if (kind == 3) t = fmin(t, sb[q]);
      else t = fmax(t, sb[q]);
    }
    *out = t;
  }
}

// ---- elementwise ----------------------------------------------------------
__global__ void k_fill(double *__restrict__ a, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_mul(double *__restrict__ a, const double *__restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = a[i] * b[i];
}
__global__ void k_inv(double *__restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 1.0 / a[i];
}
// r = f, x = 0, p = 0; dm = d*mult; bm = binv*mult (or mult when param18=1)
__global__ void k_cg_init(double *__restrict__ x, double *__restrict__ r,
                          double *__restrict__ p, const double *__restrict__ f,
                          const double *__restrict__ d, const double *__restrict__ mult,
                          const double *__restrict__ binv, double *__restrict__ dm,
                          double *__restrict__ bm, int plain_norm, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    r[i] = f[i];
    x[i] = 0.0;
    p[i] = 0.0;
    dm[i] = d[i] * mult[i];
    bm[i] = plain_norm ? mult[i] : binv[i] * mult[i];
  }
}

// x += alpha p ; r -= alpha w (when an update is pending) ; partial sums of
// (d*r)*r*mult and binv*mult*r*r (vlsc3 / vlsc32, hmholtz.f:754-759).
// VEC: 16-byte loads of node pairs (all six arrays 16-byte aligned).
template <bool VEC, bool UPD>
__global__ void __launch_bounds__(256)
    k_cg_update_dots(double *__restrict__ x, double *__restrict__ r,
                     const double *__restrict__ p, const double *__restrict__ w,
                     const double *__restrict__ dm, const double *__restrict__ bm, int64_t n,
                     const CgState *__restrict__ st, double *__restrict__ part1,
                     double *__restrict__ part2) {
  if (st->done) return;
  // UPD: the host launches the update variant for every iteration after the
  // first; a pending update exists exactly then (or the solve is done).
  constexpr bool upd = UPD;
  const double alpha = st->alpha;
  double s1 = 0.0, s2 = 0.0;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t done_upto = 0;
  if constexpr (VEC) {
    // node pairs as native 2-vectors, streamed with nontemporal hints
    typedef double v2 __attribute__((ext_vector_type(2)));
    const int64_t n2 = n >> 1;
    v2 *x2 = reinterpret_cast<v2 *>(x);
    v2 *r2 = reinterpret_cast<v2 *>(r);
    const v2 *p2 = reinterpret_cast<const v2 *>(p);
    const v2 *w2 = reinterpret_cast<const v2 *>(w);
    const v2 *dm2 = reinterpret_cast<const v2 *>(dm);
    const v2 *bm2 = reinterpret_cast<const v2 *>(bm);
    for (int64_t i = tid; i < n2; i += stride) {
      v2 ri = __builtin_nontemporal_load(r2 + i);
      if constexpr (upd) {
        v2 xi = __builtin_nontemporal_load(x2 + i);
        const v2 pi = __builtin_nontemporal_load(p2 + i), wi = __builtin_nontemporal_load(w2 + i);
        xi.x = xi.x + alpha * pi.x;
        xi.y = xi.y + alpha * pi.y;
        ri.x = ri.x + (-alpha) * wi.x;
        ri.y = ri.y + (-alpha) * wi.y;
        __builtin_nontemporal_store(xi, x2 + i);
        __builtin_nontemporal_store(ri, r2 + i);
      }
      const v2 d = __builtin_nontemporal_load(dm2 + i), b = __builtin_nontemporal_load(bm2 + i);
      const double rr0 = ri.x * ri.x, rr1 = ri.y * ri.y;
      s1 += d.x * rr0;
      s2 += b.x * rr0;
      s1 += d.y * rr1;
      s2 += b.y * rr1;
    }
    done_upto = n2 * 2;
  }
  for (int64_t i = done_upto + tid; i < n; i += stride) {
    double ri = r[i];
    if constexpr (upd) {
      x[i] = x[i] + alpha * p[i];
      ri = ri + (-alpha) * w[i];
      r[i] = ri;
    }
    const double rr = ri * ri;
    s1 += dm[i] * rr;
    s2 += bm[i] * rr;
  }
  block_store2(s1, s2, part1, part2);
}

__global__ void __launch_bounds__(kScalThreads)
    k_cg_scalars1(CgState *__restrict__ st, const double *__restrict__ part1,
                  const double *__restrict__ part2, int nparts, double *__restrict__ hist) {
  if (st->done) return;
  const double s1 = block_sum_array(part1, nparts);
  const double s2 = block_sum_array(part2, nparts);
  if (threadIdx.x != 0) return;
  st->pending = 0;
  const int iter = st->iter + 1;
  st->iter = iter;
  if (iter > st->maxit) {  // loop exhausted (hmholtz.f:817: niter = iter-1)
    st->done = 1;
    st->niter = st->maxit;
    return;
  }
  st->rtz2 = st->rtz1;
  st->rtz1 = s1;
  const double rbn2 = sqrt(s2 / st->vol);
  st->rbn2 = rbn2;
  if (iter == 1) st->rbn0 = rbn2;
  if (st->param22 < 0) st->tol = fabs(st->param22) * st->rbn0;
  if (st->tin < 0) st->tol = fabs(st->tin) * st->rbn0;
  if (hist) {
    hist[5 * (iter - 1) + 0] = rbn2;
    hist[5 * (iter - 1) + 1] = s1;
  }
  if (rbn2 <= st->tol && (iter > 1 || st->istep <= 5)) {
    st->done = 1;
    st->converged = 1;
    st->niter = iter - 1;
    return;
  }
  st->beta = (iter == 1) ? 0.0 : st->rtz1 / st->rtz2;
}

__global__ void __launch_bounds__(kScalThreads)
    k_cg_scalars2(CgState *__restrict__ st, const double *__restrict__ rp1, int n1,
                  const double *__restrict__ rp2, int n2, double *__restrict__ hist) {
  if (st->done) return;
  const double a = block_sum_array(rp1, n1);
  const double b = (n2 > 0) ? block_sum_array(rp2, n2) : 0.0;
  if (threadIdx.x != 0) return;
  const double rho = a + b;
  st->rho = rho;
  st->alpha = st->rtz1 / rho;
  st->pending = 1;
  if (hist) {
    const int it = st->iter;
    hist[5 * (it - 1) + 2] = rho;
    hist[5 * (it - 1) + 3] = st->alpha;
    hist[5 * (it - 1) + 4] = st->beta;
  }
}

// rho over all nodes (used when ids are not zero exactly at interiors):
// part[b] = sum w*p*mult over the block's nodes
__global__ void __launch_bounds__(256)
    k_dot3_guarded(const double *__restrict__ a, const double *__restrict__ b,
                   const double *__restrict__ c, int64_t n, const CgState *__restrict__ st,
                   double *__restrict__ part) {
  __shared__ double sb[4];
  if (st->done) return;
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v += a[i] * b[i] * c[i];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sb[0] + sb[1]) + sb[2]) + sb[3];
}

__global__ void k_mask_guarded(double *__restrict__ w, const double *__restrict__ mask,
                               int64_t n, const CgState *__restrict__ st) {
  if (st->done) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] * mask[i];
}

// max(|mask-1|, |mult-1|) over element-interior nodes and max |f| over
// masked nodes, one partial per block
__global__ void __launch_bounds__(256)
    k_fused_checks(const double *__restrict__ mask, const double *__restrict__ mult,
                   const double *__restrict__ f, int64_t n, int np,
                   double *__restrict__ part) {
  __shared__ double sb[4];
  const int64_t np3 = (int64_t)np * np * np;
  double v = 0.0;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(l % np3);
    const int i = q % np, j = (q / np) % np, k = q / (np * np);
    if (i > 0 && i < np - 1 && j > 0 && j < np - 1 && k > 0 && k < np - 1)
      v = fmax(v, fmax(fabs(mask[l] - 1.0), fabs(mult[l] - 1.0)));
    if (f && mask[l] == 0.0) v = fmax(v, fabs(f[l]));
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sb[0], sb[1]), fmax(sb[2], sb[3]));
}

// ---- mixed precision: FP32 iteration, FP64 solution and true residual ----
// The Jacobi PCG iterates in single precision (r, p, w, the correction xs
// accumulated since the last refresh, the preconditioner; FP32 operator),
// with every inner product accumulated in FP64.  A refresh -- x += xs in
// FP64, r = f - mask*QQ^T A_L x in FP64 (FP64 axhelm and assembly), the
// iterated residual replaced by it -- runs when the iterated residual has
// fallen by kRefresh since the last one, when it meets the tolerance (the
// solve ends only if the FP64 residual meets it too) and when maxit is
// reached.  The search direction survives a refresh (reliable updates:
// Sleijpen & van der Vorst 1996; Clark et al. 2010 for single/double CG).
struct McgState {
  CgState cg;    // the iteration (read back by the host as a CgState)
  CgState gate;  // gate.done == 0 while a refresh is due (gates its kernels)
  double rmax;   // largest iterated residual norm since the last refresh
  int final;     // 1: iterated residual met the tolerance; 2: maxit reached
  int nrefresh;
};
// Refresh when the iterated residual has fallen by this factor since the
// last refresh.  The FP32 recurrence stays within a few FP32 roundings of
// the residual it started from, so a drop of 1e4 still leaves the iterated
// residual accurate to about 1e-3 relative; a factor of 0.1 refreshed every
// second iteration of a fast solve (8 decades in 16 iterations), costing an
// FP64 operator apply each time.  Measured on five time-step solves of 10-32
// iterations: 0.1 -> 1.07x over FP64, 1e-2 -> 1.24x, 1e-3 -> 1.32x,
// 1e-4 -> 1.40x with unchanged iteration counts and FP64 residuals; 1e-5
// and 1e-6 gain nothing more and let the solution drift further (up to
// 9e-9 relative from the FP64 solve at 1e-6, 4e-11 at 1e-4).
constexpr double kRefresh = 1e-4;
// grid of the refresh kernels: they run every pass of a captured loop and
// mostly return at once, at a cost that grows with the number of blocks
// (about 10 ns each), so it scales with the field: 64Ki nodes per block,
// 64 to 2048 blocks (a multiple of 8, as the assembly's XCD mapping needs)
constexpr int kRefreshBlocks = 2048;
inline int refresh_blocks(int64_t n) {
  int64_t b = (n + 65535) / 65536;
  b = (b + 7) / 8 * 8;
  return (int)(b < 64 ? 64 : (b > kRefreshBlocks ? kRefreshBlocks : b));
}

// xs += alpha p ; r -= alpha w (FP32) ; FP64 partials of dm*r*r and bm*r*r
template <bool VEC, bool UPD>
__global__ void __launch_bounds__(256)
    k_mcg_update_dots(float *__restrict__ xs, float *__restrict__ r,
                      const float *__restrict__ p, const float *__restrict__ w,
                      const float *__restrict__ dm, const float *__restrict__ bm, int64_t n,
                      const CgState *__restrict__ st, double *__restrict__ part1,
                      double *__restrict__ part2) {
  if (st->done) return;
  const float alpha = (float)st->alpha;
  double s1 = 0.0, s2 = 0.0;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t done_upto = 0;
  if constexpr (VEC) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const int64_t n4 = n >> 2;
    v4 *x4 = reinterpret_cast<v4 *>(xs);
    v4 *r4 = reinterpret_cast<v4 *>(r);
    const v4 *p4 = reinterpret_cast<const v4 *>(p);
    const v4 *w4 = reinterpret_cast<const v4 *>(w);
    const v4 *d4 = reinterpret_cast<const v4 *>(dm);
    const v4 *b4 = reinterpret_cast<const v4 *>(bm);
    for (int64_t i = tid; i < n4; i += stride) {
      v4 ri = __builtin_nontemporal_load(r4 + i);
      if constexpr (UPD) {
        v4 xi = __builtin_nontemporal_load(x4 + i);
        const v4 pi = __builtin_nontemporal_load(p4 + i), wi = __builtin_nontemporal_load(w4 + i);
        xi.x += alpha * pi.x;
        xi.y += alpha * pi.y;
        xi.z += alpha * pi.z;
        xi.w += alpha * pi.w;
        ri.x -= alpha * wi.x;
        ri.y -= alpha * wi.y;
        ri.z -= alpha * wi.z;
        ri.w -= alpha * wi.w;
        __builtin_nontemporal_store(xi, x4 + i);
        __builtin_nontemporal_store(ri, r4 + i);
      }
      const v4 d = __builtin_nontemporal_load(d4 + i), b = __builtin_nontemporal_load(b4 + i);
      const double q0 = (double)ri.x * ri.x, q1 = (double)ri.y * ri.y,
                   q2 = (double)ri.z * ri.z, q3 = (double)ri.w * ri.w;
      s1 += d.x * q0 + d.y * q1 + d.z * q2 + d.w * q3;
      s2 += b.x * q0 + b.y * q1 + b.z * q2 + b.w * q3;
    }
    done_upto = n4 * 4;
  }
  for (int64_t i = done_upto + tid; i < n; i += stride) {
    float ri = r[i];
    if constexpr (UPD) {
      xs[i] = xs[i] + alpha * p[i];
      ri = ri - alpha * w[i];
      r[i] = ri;
    }
    const double q = (double)ri * ri;
    s1 += dm[i] * q;
    s2 += bm[i] * q;
  }
  block_store2(s1, s2, part1, part2);
}

// as k_cg_scalars1, but a refresh of the residual is requested instead of
// ending the solve or when the iterated residual fell by kRefresh
__global__ void __launch_bounds__(kScalThreads)
    k_mcg_scalars1(McgState *__restrict__ ms, const double *__restrict__ part1,
                   const double *__restrict__ part2, int nparts, double *__restrict__ hist) {
  CgState *st = &ms->cg;
  if (st->done) return;
  const double s1 = block_sum_array(part1, nparts);
  const double s2 = block_sum_array(part2, nparts);
  if (threadIdx.x != 0) return;
  const int iter = st->iter + 1;
  st->iter = iter;
  if (iter > st->maxit) {  // loop exhausted: fold the correction into x
    ms->final = 2;
    ms->gate.done = 0;
    return;
  }
  st->rtz2 = st->rtz1;
  st->rtz1 = s1;
  const double rbn2 = sqrt(s2 / st->vol);
  st->rbn2 = rbn2;
  if (iter == 1) {
    st->rbn0 = rbn2;
    ms->rmax = rbn2;
  }
  if (st->param22 < 0) st->tol = fabs(st->param22) * st->rbn0;
  if (st->tin < 0) st->tol = fabs(st->tin) * st->rbn0;
  if (hist) {
    hist[5 * (iter - 1) + 0] = rbn2;
    hist[5 * (iter - 1) + 1] = s1;
  }
  if (rbn2 <= st->tol && (iter > 1 || st->istep <= 5)) {
    ms->final = 1;
    ms->gate.done = 0;
    return;
  }
  if (rbn2 < kRefresh * ms->rmax) {
    ms->final = 0;
    ms->gate.done = 0;
    return;
  }
  if (rbn2 > ms->rmax) ms->rmax = rbn2;
  st->beta = (iter == 1) ? 0.0 : st->rtz1 / st->rtz2;
}

// refresh, part 1: x += xs (FP64), xs = 0
__global__ void k_mcg_fold(double *__restrict__ x, float *__restrict__ xs, int64_t n,
                           const CgState *__restrict__ gate) {
  if (gate->done) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = x[i] + (double)xs[i];
    xs[i] = 0.0f;
  }
}

// refresh, part 2 (after ax = mask*QQ^T A_L x): r = f - ax in FP64, stored
// rounded; partials of dm*r32^2 (the iteration's rtz) and bm*r^2 (FP64 norm)
__global__ void __launch_bounds__(256)
    k_mcg_resid(float *__restrict__ r32, const double *__restrict__ f,
                const double *__restrict__ ax, const float *__restrict__ dm,
                const float *__restrict__ bm, int64_t n, const CgState *__restrict__ gate,
                double *__restrict__ part1, double *__restrict__ part2) {
  if (gate->done) return;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = f[i] - ax[i];
    const float rf = (float)ri;
    r32[i] = rf;
    s1 += dm[i] * ((double)rf * rf);
    s2 += bm[i] * (ri * ri);
  }
  block_store2(s1, s2, part1, part2);
}

// refresh, part 3: the true residual's norm decides the exit; otherwise the
// iteration continues from it with the same search direction
__global__ void __launch_bounds__(kScalThreads)
    k_mcg_scalars1b(McgState *__restrict__ ms, const double *__restrict__ part1,
                    const double *__restrict__ part2, int nparts, double *__restrict__ hist) {
  if (ms->gate.done) return;
  const double s1 = block_sum_array(part1, nparts);
  const double s2 = block_sum_array(part2, nparts);
  if (threadIdx.x != 0) return;
  CgState *st = &ms->cg;
  ms->gate.done = 1;
  ms->nrefresh += 1;
  const double rbn2 = sqrt(s2 / st->vol);
  st->rbn2 = rbn2;
  if (ms->final == 2) {
    st->done = 1;
    st->niter = st->maxit;
    return;
  }
  const int iter = st->iter;
  st->rtz1 = s1;
  if (hist) {
    hist[5 * (iter - 1) + 0] = rbn2;
    hist[5 * (iter - 1) + 1] = s1;
  }
  if (ms->final == 1 && rbn2 <= st->tol) {
    st->done = 1;
    st->converged = 1;
    st->niter = iter - 1;
    return;
  }
  ms->final = 0;
  ms->rmax = rbn2;
  st->beta = (iter == 1) ? 0.0 : st->rtz1 / st->rtz2;
}

// x = 0, xs = 0, p = 0, r = f (rounded), d, d*mult and binv*mult (or mult)
// in FP32
__global__ void k_mcg_init(double *__restrict__ x, float *__restrict__ xs,
                           float *__restrict__ p, float *__restrict__ r,
                           float *__restrict__ d32, float *__restrict__ dm,
                           float *__restrict__ bm, const double *__restrict__ f,
                           const double *__restrict__ d, const double *__restrict__ mult,
                           const double *__restrict__ binv, int plain_norm, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = 0.0;
    xs[i] = 0.0f;
    p[i] = 0.0f;
    r[i] = (float)f[i];
    d32[i] = (float)d[i];
    dm[i] = (float)(d[i] * mult[i]);
    bm[i] = (float)(plain_norm ? mult[i] : binv[i] * mult[i]);
  }
}

int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

// Reduce on the device and read the scalar back (synchronous).
int reduce_host(const Mesh &m, const double *a, const double *b, const double *c, int64_t n,
                int kind, double *out, hipStream_t s) {
  NK_CHECK(m.red != nullptr, NKH_ERR_STATE, "reduction buffer not allocated");
  hipLaunchKernelGGL(k_reduce, dim3(kRedBlocks), dim3(256), 0, s, a, b, c, n, kind,
                     m.red + kRedScratch);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, m.red + kRedScratch, kRedBlocks,
                     kind, m.red + kRedOut);
  NK_HIP(hipGetLastError());
  NK_HIP(hipMemcpyAsync(out, m.red + kRedOut, sizeof(double), hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

}  // namespace

int ensure_small(Mesh &m) {
  // room for the mixed-precision solve's state (a CgState first)
  if (!m.st) NK_HIP(hipMalloc(&m.st, sizeof(McgState)));
  if (!m.red) NK_HIP(hipMalloc(&m.red, sizeof(double) * kRedSize));
  return NKH_OK;
}

int ensure_workspace(Mesh &m) {
  NK_TRY(ensure_small(m));
  const size_t need = cg_ws_doubles(m.n);
  if (m.ws && m.ws_doubles >= need) return NKH_OK;
  if (m.ws) (void)hipFree(m.ws);
  m.ws = nullptr;
  if (hipMalloc(&m.ws, sizeof(double) * need) != hipSuccess) {
    set_error("workspace allocation of %zu doubles failed", need);
    return NKH_ERR_NOMEM;
  }
  m.ws_doubles = need;
  return NKH_OK;
}

// Reduce a coefficient field to a uniform value when it is constant (the
// multiply by the scalar is then bit-identical to the multiply by the field).
static int classify(Mesh &m, const double *f, Coef *c, hipStream_t s) {
  NK_TRY(ensure_small(m));
  double mn, mx;
  NK_TRY(reduce_host(m, f, nullptr, nullptr, m.n, 3, &mn, s));
  NK_TRY(reduce_host(m, f, nullptr, nullptr, m.n, 2, &mx, s));
  if (mn == mx) {
    c->f = nullptr;
    c->s = mn;
  } else {
    c->f = f;
    c->s = 0.0;
  }
  return NKH_OK;
}

int classify_coefs(Mesh &m, const double *h1, const double *h2, Coef *c1, Coef *c2,
                   hipStream_t s) {
  NK_TRY(classify(m, h1, c1, s));
  // setfast (hmholtz.f:303-305): ifh2 = vlamax(h2) > 0; a uniform 0 drops the term
  NK_TRY(classify(m, h2, c2, s));
  return NKH_OK;
}

int mul_inplace(double *a, const double *b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int run_setprec(Mesh &m, double *d, Coef h1, Coef h2, hipStream_t s) {
  NK_TRY(launch_setprec(m, d, h1, h2, s));
  NK_TRY(gs_op(m.gs, d, 0, s));
  hipLaunchKernelGGL(k_inv, dim3(grid_for(m.n)), dim3(256), 0, s, d, m.n);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// The conditions under which a CG denominator (mask * QQ^T w, p)_mult may be
// taken as sum_l p_l w_l before the assembly: ids zero exactly at element
// interiors, mask 1 and mult 1 there, f (when given) zero where masked, and
// per group a 0/1 mask and mult summing to 1.  Leaves the group flags of
// gs_prepare in m.gs.
int fused_conditions(Mesh &m, const double *mask, const double *mult, const double *f, bool *ok,
                     hipStream_t s) {
  *ok = false;
  if (!m.gs.interior_zero) return NKH_OK;
  NK_TRY(ensure_small(m));
  hipLaunchKernelGGL(k_fused_checks, dim3(kRedBlocks), dim3(256), 0, s, mask, mult, f, m.n, m.np,
                     m.red + kRedScratch);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, m.red + kRedScratch, kRedBlocks,
                     2, m.red + kRedOut);
  NK_HIP(hipGetLastError());
  double dev;
  NK_HIP(hipMemcpyAsync(&dev, m.red + kRedOut, sizeof(double), hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  if (dev != 0.0) return NKH_OK;
  return gs_prepare(m.gs, mask, mult, m.red + kRedScratch, ok, s);
}

int run_chktcg1(Mesh &m, double *tol, const double *res, Coef h1, Coef h2,
                const double *mask, const double *mult) {
  hipStream_t s = cur_stream();
  NK_TRY(ensure_workspace(m));
  const int64_t n = m.n;
  double *w1 = m.ws, *w2 = m.ws + n;
  const double eps = 1.0e-13, acondno = 10.0;
  double a;
  // rinit = sqrt(glsc3(binv*res, res, mult)/vol)
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, w1, 1.0, n);
  NK_HIP(hipMemcpyAsync(w2, m.binv, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, s, w2, res, n);
  NK_TRY(reduce_host(m, w2, res, mult, n, 0, &a, s));
  const double rinit = std::sqrt(a / m.vol);
  const double rmin = eps * rinit;
  if (*tol < rmin) *tol = rmin;
  double bcneu1, bcneu2;
  NK_TRY(reduce_host(m, w1, mask, mult, n, 0, &bcneu1, s));
  NK_TRY(reduce_host(m, w1, w1, mult, n, 0, &bcneu2, s));
  const double bctest = std::fabs(bcneu1 - bcneu2);
  NK_TRY(launch_axhelm(m, w2, w1, h1, h2, s));
  double sq;
  NK_TRY(reduce_host(m, w2, w2, m.bm1, n, 0, &sq, s));
  const double bcrob = std::sqrt(sq / m.vol);
  if (bctest < 0.1 && bcrob < eps * acondno) {
    const double tolmin = rinit * eps * 10.0;
    if (*tol < tolmin) *tol = tolmin;
  }
  return NKH_OK;
}

// One Jacobi-PCG solve whose iterations can be enqueued one at a time (the
// single solve below, and the batch of solves captured in a graph).
struct CgRun {
  Mesh *m = nullptr;
  double *x = nullptr, *r = nullptr, *w = nullptr, *p = nullptr, *d = nullptr, *dm = nullptr,
         *bm = nullptr;
  double *red = nullptr;
  CgState *st = nullptr;
  const uint8_t *flag = nullptr;  // gs group flags for this solve's mask/mult
  Coef h1, h2;
  const double *mask = nullptr, *mult = nullptr;
  bool fused = false, vec_ok = false, trivial = false;
  int niter = 0, launched = 0, ax_grid_parts = 0;
  double *hist = nullptr;

  // hmholtz.f:611-720 up to the loop: preconditioner, r = f, null-space check,
  // choice of the fused iteration, device state.  ws: 6 n doubles, red:
  // kRedSize doubles, st: the state; all on stream s (host waits).
  int begin(Mesh &mesh, double *ws, double *red_, CgState *st_, double *x_, const double *f,
            Coef h1_, Coef h2_, const double *mask_, const double *mult_, double tin, int maxit,
            const nkh_params &prm, bool want_hist, hipStream_t s) {
    m = &mesh;
    const int64_t n = mesh.n;
    x = x_;
    r = ws;
    w = ws + n;
    p = ws + 2 * n;
    d = ws + 3 * n;
    dm = ws + 4 * n;
    bm = ws + 5 * n;
    red = red_;
    st = st_;
    h1 = h1_;
    h2 = h2_;
    mask = mask_;
    mult = mult_;
    double tol = std::fabs(tin);
    if (prm.restol != 0.0) tol = prm.restol;
    if (tin < 0) tol = std::fabs(tin);
    niter = maxit < kMaxCg ? maxit : kMaxCg;
    NK_TRY(run_setprec(mesh, d, h1, h2, s));
    hipLaunchKernelGGL(k_cg_init, dim3(grid_for(n)), dim3(256), 0, s, x, r, p, f, d, mult,
                       mesh.binv, dm, bm, prm.param18 == 1 ? 1 : 0, n);
    NK_HIP(hipGetLastError());
    double fmax;
    NK_TRY(reduce_host(mesh, f, nullptr, nullptr, n, 1, &fmax, s));
    trivial = fmax == 0.0;
    if (trivial) return NKH_OK;
    // null-space check (hmholtz.f:705-720): pure Neumann with h2 == 0
    double h2max, skmin;
    if (h2.uniform()) h2max = h2.s;
    else NK_TRY(reduce_host(mesh, h2.f, nullptr, nullptr, n, 2, &h2max, s));
    NK_TRY(reduce_host(mesh, mask, nullptr, nullptr, n, 3, &skmin, s));
    NK_CHECK(!(skmin > 0 && h2max == 0), NKH_ERR_UNSUPPORTED,
             "cggo: pure-Neumann mean correction (ifmcor) is not implemented");
    // Fused iteration (rho inside axhelm, mask inside the assembly) when its
    // conditions hold: ids zero exactly at element interiors, mask 1 and
    // mult 1 there, f zero where masked, and per group a 0/1 mask and mult
    // summing to 1.  Checked once per solve.
    NK_TRY(fused_conditions(mesh, mask, mult, f, &fused, s));
    flag = mesh.gs.flag;
    ax_grid_parts = axhelm_cg_grid(mesh);
    auto al16 = [](const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
    vec_ok = al16(x) && al16(r) && al16(p) && al16(w) && al16(dm) && al16(bm);
    CgState h{};
    h.rtz1 = 1.0;
    h.tol = tol;
    h.tin = tin;
    h.vol = mesh.vol;
    h.param22 = prm.param22;
    h.maxit = niter;
    h.istep = prm.istep;
    h.param18 = prm.param18;
    NK_HIP(hipMemcpyAsync(st, &h, sizeof(CgState), hipMemcpyHostToDevice, s));
    hist = nullptr;
    if (want_hist) {
      hist = red + kRedHist;
      hipLaunchKernelGGL(k_fill, dim3(grid_for(5 * kMaxCg)), dim3(256), 0, s, hist, 0.0,
                         (int64_t)5 * kMaxCg);
    }
    NK_HIP(hipGetLastError());
    return NKH_OK;
  }

  // one iteration's launches on s; ev (optional, 3 events) brackets axhelm
  // and the assembly
  int enqueue(hipStream_t s, hipEvent_t *ev = nullptr) {
    const int64_t n = m->n;
    double *part1 = red + kRedPart1, *part2 = red + kRedPart2, *rho1 = red + kRedRho1,
           *rho2 = red + kRedRho2;
    auto kern = vec_ok ? (launched == 0 ? k_cg_update_dots<true, false>
                                        : k_cg_update_dots<true, true>)
                       : (launched == 0 ? k_cg_update_dots<false, false>
                                        : k_cg_update_dots<false, true>);
    const int ub = upd_blocks(n);
    hipLaunchKernelGGL(kern, dim3(ub), dim3(256), 0, s, x, r, p, w, dm, bm, n, st, part1, part2);
    hipLaunchKernelGGL(k_cg_scalars1, dim3(1), dim3(kScalThreads), 0, s, st, part1, part2, ub,
                       hist);
    if (fused) {
      if (ev) NK_HIP(hipEventRecord(ev[0], s));
      NK_TRY(launch_axhelm_cg(*m, w, p, r, d, h1, h2, mult, st, rho1, s));
      if (ev) NK_HIP(hipEventRecord(ev[1], s));
      NK_TRY(gs_assemble(m->gs, w, st, s, flag));
      if (ev) NK_HIP(hipEventRecord(ev[2], s));
      hipLaunchKernelGGL(k_cg_scalars2, dim3(1), dim3(kScalThreads), 0, s, st, rho1,
                         ax_grid_parts, rho2, 0, hist);
    } else {
      NK_TRY(launch_axhelm_cg(*m, w, p, r, d, h1, h2, mult, st, nullptr, s));
      NK_TRY(gs_op(m->gs, w, 0, s));
      hipLaunchKernelGGL(k_mask_guarded, dim3(grid_for(n)), dim3(256), 0, s, w, mask, n, st);
      hipLaunchKernelGGL(k_dot3_guarded, dim3(kRedBlocks), dim3(256), 0, s, w, p, mult, n, st,
                         rho1);
      hipLaunchKernelGGL(k_cg_scalars2, dim3(1), dim3(kScalThreads), 0, s, st, rho1, kRedBlocks,
                         rho2, 0, hist);
    }
    NK_HIP(hipGetLastError());
    ++launched;
    return NKH_OK;
  }

  void result(const CgState &hs, nkh_result *res) const {
    nkh_result out{};
    if (!trivial) {
      out.niter = hs.niter;
      out.converged = hs.converged;
      out.rbn0 = hs.rbn0;
      out.rbn2 = hs.rbn2;
      out.tol = hs.tol;
    }
    if (res) *res = out;
  }
};

// The mixed-precision Jacobi PCG (see McgState), enqueued like CgRun.
// Work space (cg_ws_doubles): ax (n doubles, FP64 refresh and set-up
// scratch), then seven FP32 vectors r, p, w, xs, d, d*mult, bm at 256-byte
// aligned offsets.  Needs uniform h1/h2 and the fused iteration's conditions.
struct McgRun {
  Mesh *m = nullptr;
  double *x = nullptr, *ax = nullptr;
  const double *f = nullptr;
  float *r = nullptr, *p = nullptr, *w = nullptr, *xs = nullptr, *d = nullptr, *dm = nullptr,
        *bm = nullptr;
  double *red = nullptr;
  McgState *ms = nullptr;
  CgState *st = nullptr;  // &ms->cg
  const uint8_t *flag = nullptr;
  Coef h1, h2;
  bool trivial = false, vec_ok = false, fused = true;
  int niter = 0, launched = 0, ax_grid_parts = 0;
  double *hist = nullptr;

  int begin(Mesh &mesh, double *ws, double *red_, void *st_, double *x_, const double *f_,
            Coef h1_, Coef h2_, const double *mask, const double *mult, double tin, int maxit,
            const nkh_params &prm, bool want_hist, hipStream_t s) {
    m = &mesh;
    const int64_t n = mesh.n;
    NK_CHECK(h1_.uniform() && h2_.uniform(), NKH_ERR_UNSUPPORTED,
             "mixed-precision cggo: h1 and h2 must be uniform");
    x = x_;
    f = f_;
    ax = ws;
    const int64_t nn = (n + 63) / 64 * 64;
    float *base = reinterpret_cast<float *>(ws + (n + 31) / 32 * 32);
    r = base;
    p = base + nn;
    w = base + 2 * nn;
    xs = base + 3 * nn;
    d = base + 4 * nn;
    dm = base + 5 * nn;
    bm = base + 6 * nn;
    red = red_;
    ms = static_cast<McgState *>(st_);
    st = &ms->cg;
    h1 = h1_;
    h2 = h2_;
    double tol = std::fabs(tin);
    if (prm.restol != 0.0) tol = prm.restol;
    if (tin < 0) tol = std::fabs(tin);
    niter = maxit < kMaxCg ? maxit : kMaxCg;
    NK_TRY(run_setprec(mesh, ax, h1, h2, s));  // d (FP64) in the ax slot
    hipLaunchKernelGGL(k_mcg_init, dim3(grid_for(n)), dim3(256), 0, s, x, xs, p, r, d, dm, bm, f,
                       ax, mult, mesh.binv, prm.param18 == 1 ? 1 : 0, n);
    NK_HIP(hipGetLastError());
    double fmax;
    NK_TRY(reduce_host(mesh, f, nullptr, nullptr, n, 1, &fmax, s));
    trivial = fmax == 0.0;
    if (trivial) return NKH_OK;
    double skmin;
    NK_TRY(reduce_host(mesh, mask, nullptr, nullptr, n, 3, &skmin, s));
    NK_CHECK(!(skmin > 0 && h2.s == 0.0), NKH_ERR_UNSUPPORTED,
             "cggo: pure-Neumann mean correction (ifmcor) is not implemented");
    bool ok = false;
    NK_TRY(fused_conditions(mesh, mask, mult, f, &ok, s));
    NK_CHECK(ok, NKH_ERR_UNSUPPORTED,
             "mixed-precision cggo: needs ids zero at element interiors, 0/1 masks and "
             "mult summing to 1 per group");
    flag = mesh.gs.flag;
    NK_TRY(ensure_geom32(mesh, s));
    ax_grid_parts = axhelm_cg_grid(mesh);
    auto al16 = [](const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
    vec_ok = al16(r) && al16(p) && al16(w) && al16(xs) && al16(dm) && al16(bm);
    McgState h{};
    h.cg.rtz1 = 1.0;
    h.cg.tol = tol;
    h.cg.tin = tin;
    h.cg.vol = mesh.vol;
    h.cg.param22 = prm.param22;
    h.cg.maxit = niter;
    h.cg.istep = prm.istep;
    h.cg.param18 = prm.param18;
    h.gate.done = 1;
    NK_HIP(hipMemcpyAsync(ms, &h, sizeof(McgState), hipMemcpyHostToDevice, s));
    hist = nullptr;
    if (want_hist) {
      hist = red + kRedHist;
      hipLaunchKernelGGL(k_fill, dim3(grid_for(5 * kMaxCg)), dim3(256), 0, s, hist, 0.0,
                         (int64_t)5 * kMaxCg);
    }
    NK_HIP(hipGetLastError());
    return NKH_OK;
  }

  int enqueue(hipStream_t s, hipEvent_t *ev = nullptr) {
    const int64_t n = m->n;
    double *part1 = red + kRedPart1, *part2 = red + kRedPart2, *rho1 = red + kRedRho1,
           *rho2 = red + kRedRho2;
    auto kern = vec_ok ? (launched == 0 ? k_mcg_update_dots<true, false>
                                        : k_mcg_update_dots<true, true>)
                       : (launched == 0 ? k_mcg_update_dots<false, false>
                                        : k_mcg_update_dots<false, true>);
    const int ub = upd_blocks(n);
    hipLaunchKernelGGL(kern, dim3(ub), dim3(256), 0, s, xs, r, p, w, dm, bm, n, st, part1, part2);
    hipLaunchKernelGGL(k_mcg_scalars1, dim3(1), dim3(kScalThreads), 0, s, ms, part1, part2, ub,
                       hist);
    // the refresh (its kernels return at once while the gate is closed)
    const CgState *gate = &ms->gate;
    const int rb = refresh_blocks(n);
    hipLaunchKernelGGL(k_mcg_fold, dim3(rb), dim3(256), 0, s, x, xs, n, gate);
    NK_TRY(launch_axhelm_gated(*m, ax, x, h1, h2, gate, s));
    NK_TRY(gs_assemble(m->gs, ax, gate, s, flag, rb));
    hipLaunchKernelGGL(k_mcg_resid, dim3(rb), dim3(256), 0, s, r, f, ax, dm, bm, n, gate, part1,
                       part2);
    hipLaunchKernelGGL(k_mcg_scalars1b, dim3(1), dim3(kScalThreads), 0, s, ms, part1, part2, rb,
                       hist);
    if (ev) NK_HIP(hipEventRecord(ev[0], s));
    NK_TRY(launch_axhelm32_jac(*m, w, p, r, d, (float)h1.s, (float)h2.s, st, rho1, s));
    if (ev) NK_HIP(hipEventRecord(ev[1], s));
    NK_TRY(gs_assemble(m->gs, w, st, s, flag));
    if (ev) NK_HIP(hipEventRecord(ev[2], s));
    hipLaunchKernelGGL(k_cg_scalars2, dim3(1), dim3(kScalThreads), 0, s, st, rho1,
                       ax_grid_parts, rho2, 0, hist);
    NK_HIP(hipGetLastError());
    ++launched;
    return NKH_OK;
  }

  void result(const CgState &hs, nkh_result *res) const {
    nkh_result out{};
    if (!trivial) {
      out.niter = hs.niter;
      out.converged = hs.converged;
      out.rbn0 = hs.rbn0;
      out.rbn2 = hs.rbn2;
      out.tol = hs.tol;
    }
    if (res) *res = out;
  }
};

size_t cg_ws_doubles(int64_t n) {
  const size_t mixed = (size_t)((n + 31) / 32 * 32) + (size_t)(7 * ((n + 63) / 64 * 64) + 1) / 2;
  return std::max((size_t)6 * (size_t)n, mixed);
}

namespace {

template <class Run>
int drive_cggo(Mesh &m, double *x, const double *f, Coef h1, Coef h2, const double *mask,
               const double *mult, double tin, int maxit, const nkh_params &prm,
               nkh_result *res, double *history_host, bool allow_timers = true) {
  hipStream_t s = cur_stream();
  NK_TRY(ensure_workspace(m));
  Run run;
  NK_TRY(run.begin(m, m.ws, m.red, m.st, x, f, h1, h2, mask, mult, tin, maxit, prm,
                   history_host != nullptr, s));
  if (run.trivial) {
    run.result(CgState{}, res);
    return NKH_OK;
  }
  CgState hs{};
  int check_every = 1;
  // CTIMER mirror: event pairs around axhelm and gs of every launched iteration
  const bool timed = allow_timers && timers().on;
  const int niter = run.niter;
  std::vector<hipEvent_t> ev;
  if (timed) {
    ev.resize((size_t)4 * (niter + 1));
    for (auto &e : ev) NK_HIP(hipEventCreate(&e));
  }
  hipEvent_t t_solve0 = nullptr, t_solve1 = nullptr;
  if (timed) {
    NK_HIP(hipEventCreate(&t_solve0));
    NK_HIP(hipEventCreate(&t_solve1));
    NK_HIP(hipEventRecord(t_solve0, s));
  }
  while (true) {
    NK_TRY(run.enqueue(s, timed ? &ev[4 * run.launched] : nullptr));
    if (run.launched % check_every == 0 || run.launched > niter) {
      NK_HIP(hipMemcpyAsync(&hs, run.st, sizeof(CgState), hipMemcpyDeviceToHost, s));
      NK_HIP(hipStreamSynchronize(s));
      if (hs.done) break;
      NK_CHECK(run.launched <= niter + 1, NKH_ERR_STATE, "cggo: state machine did not terminate");
      if (run.launched >= 8) check_every = 4;
    }
  }
  if (timed) {
    NK_HIP(hipEventRecord(t_solve1, s));
    NK_HIP(hipEventSynchronize(t_solve1));
    float ms = 0.f;
    // axhelm and gs did work in launches 1..niter (later ones returned at once)
    for (int it = 0; it < hs.niter && run.fused; ++it) {
      NK_HIP(hipEventElapsedTime(&ms, ev[4 * it + 0], ev[4 * it + 1]));
      timers().taxhm += ms * 1e-3;
      timers().naxhm += 1;
      NK_HIP(hipEventElapsedTime(&ms, ev[4 * it + 1], ev[4 * it + 2]));
      timers().tdsum += ms * 1e-3;
      timers().ndsum += 1;
    }
    NK_HIP(hipEventElapsedTime(&ms, t_solve0, t_solve1));
    timers().thmhz += ms * 1e-3;
    timers().nhmhz += 1;
    for (auto &e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t_solve0);
    (void)hipEventDestroy(t_solve1);
  }
  run.result(hs, res);
  if (history_host) {
    NK_HIP(hipMemcpyAsync(history_host, run.hist, sizeof(double) * 5 * kMaxCg,
                          hipMemcpyDeviceToHost, s));
    NK_HIP(hipStreamSynchronize(s));
  }
  return NKH_OK;
}

}  // namespace

int run_cggo(Mesh &m, double *x, const double *f, Coef h1, Coef h2, const double *mask,
             const double *mult, double tin, int maxit, const nkh_params &prm,
             nkh_result *res, double *history_host) {
  if (prm.precision == 2)
    return drive_cggo<McgRun>(m, x, f, h1, h2, mask, mult, tin, maxit, prm, res, history_host);
  return drive_cggo<CgRun>(m, x, f, h1, h2, mask, mult, tin, maxit, prm, res, history_host);
}

// ---------------------------------------------------------------------------
// Allocation-free FP64 Jacobi PCG (nkd_cggo_noalloc), for callers that may
// not allocate device memory.  The solve's vectors (r, w, p, d, d*mult,
// bm), its partial sums and history, its device state and the gs group
// flags of its mask/mult live in `scratch`, one file-scope device array per
// GPU, sized for meshes of at most kNoAllocMaxN local nodes.  Host code
// reaches the array only through the address hipGetSymbolAddress returns
// for the current device (noalloc_scratch).  Every call on a GPU uses the
// same array: calls that overlap on one GPU race.  Nothing in it needs to
// start at zero (each solve writes its state, partial sums, flags and
// vectors before reading them); the last call's contents stay behind.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocWs = 6 * (size_t)kNoAllocMaxN;                         // doubles
constexpr size_t kNoAllocRed = ((size_t)kRedSize + 31) / 32 * 32;               // doubles
constexpr size_t kNoAllocState = (sizeof(McgState) + 255) / 256 * 32;           // doubles
constexpr size_t kNoAllocFlag = ((size_t)kNoAllocMaxN + 255) / 256 * 32;        // doubles
constexpr size_t kNoAllocDoubles = kNoAllocWs + kNoAllocRed + kNoAllocState + kNoAllocFlag;
static_assert(kNoAllocMaxN < ((int64_t)1 << 20),
              "the axhelm autotuner (which allocates) runs on meshes of 2^20 nodes and up");

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kNoAllocDevices = 64;
double *g_scratch_dev[kNoAllocDevices] = {};
}  // namespace

int noalloc_scratch(double **out) {
  int dev = 0;
  NK_HIP(hipGetDevice(&dev));
  NK_CHECK(dev >= 0 && dev < kNoAllocDevices, NKH_ERR_ARG, "noalloc scratch: device %d", dev);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    NK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NK_CHECK(p != nullptr, NKH_ERR_HIP, "noalloc scratch: no device address");
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  if (out) *out = g_scratch_dev[dev];
  return NKH_OK;
}

int run_cggo_noalloc(Mesh &m, double *x, const double *f, const double *h1, const double *h2,
                     const double *mask, const double *mult, double tin, int maxit,
                     const nkh_params &prm, nkh_result *res, double *history_host) {
  NK_CHECK(m.n <= kNoAllocMaxN, NKH_ERR_ARG,
           "nkd_cggo_noalloc: %lld local nodes exceed its static scratch (at most %lld)",
           (long long)m.n, (long long)kNoAllocMaxN);
  NK_CHECK(m.gs.ngroups <= kNoAllocMaxN, NKH_ERR_ARG, "nkd_cggo_noalloc: %lld gs groups",
           (long long)m.gs.ngroups);
  double *sp = nullptr;
  NK_TRY(noalloc_scratch(&sp));
  // a view of the mesh whose solve buffers are the scratch: the geometry
  // and gather-scatter map are the mesh's own, nothing of it is changed
  Mesh v = m;
  v.ws = sp;
  v.ws_doubles = kNoAllocWs;
  v.red = sp + kNoAllocWs;
  v.st = reinterpret_cast<CgState *>(sp + kNoAllocWs + kNoAllocRed);
  v.gs.flag = reinterpret_cast<uint8_t *>(sp + kNoAllocWs + kNoAllocRed + kNoAllocState);
  Coef c1, c2;
  NK_TRY(classify_coefs(v, h1, h2, &c1, &c2, cur_stream()));
  return drive_cggo<CgRun>(v, x, f, c1, c2, mask, mult, tin, maxit, prm, res, history_host,
                           false);
}

// ---------------------------------------------------------------------------
// Batched Jacobi-PCG: several hmholtz solves (different meshes, h1/h2,
// masks, sizes; the same lx1 or not) iterate together.  Each solve keeps its
// own work space, state and gs group flags; kGraphIters steady-state
// iterations of all of them -- each solve's launches on its own stream,
// forked from and joined to a library stream once per replay -- are
// captured once in a hipGraph and replayed until every solve has stopped (a
// finished solve's kernels return at once).  No memset is recorded in the
// graph.
// ---------------------------------------------------------------------------
// iterations per replay of the batch graph (the host reads the states after
// each replay)
constexpr int kGraphIters = 8;

struct CgBatch {
  int nb = 0;
  std::vector<int> mesh;
  std::vector<double *> ws, red;
  std::vector<CgState *> st;
  std::vector<uint8_t *> flag;
  std::vector<int64_t> nflag;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> ev;  // fork + one join per solve
  // the graph is captured and launched on a library stream (capture is not
  // allowed on the caller's stream when that is the null stream), ordered
  // after and before the caller's stream by ev_sync
  hipStream_t cap = nullptr;
  hipEvent_t ev_sync = nullptr;
  void release() {
    if (cap) (void)hipStreamDestroy(cap);
    if (ev_sync) (void)hipEventDestroy(ev_sync);
    cap = nullptr;
    ev_sync = nullptr;
    for (double *q : ws) (void)hipFree(q);
    for (double *q : red) (void)hipFree(q);
    for (CgState *q : st) (void)hipFree(q);
    for (uint8_t *q : flag) (void)hipFree(q);
    for (hipStream_t q : streams) (void)hipStreamDestroy(q);
    for (hipEvent_t q : ev) (void)hipEventDestroy(q);
    ws.clear();
    red.clear();
    st.clear();
    flag.clear();
    streams.clear();
    ev.clear();
  }
};

namespace {
std::vector<CgBatch *> g_batches;
}

int batch_create(int nb, const int *meshes, int *id) {
  NK_CHECK(nb >= 1 && nb <= 64 && meshes && id, NKH_ERR_ARG, "nkh_batch_create: bad arguments");
  CgBatch *B = new CgBatch();
  B->nb = nb;
  auto fail = [&](int rc) {
    B->release();
    delete B;
    return rc;
  };
  for (int k = 0; k < nb; ++k) {
    Mesh *m = get_mesh(meshes[k]);
    if (!m || !m->has_geom || !m->has_gs) {
      set_error("nkh_batch_create: mesh %d is not ready", meshes[k]);
      return fail(NKH_ERR_STATE);
    }
    NK_TRY(ensure_small(*m));
    double *w = nullptr, *r = nullptr;
    CgState *c = nullptr;
    uint8_t *fl = nullptr;
    hipStream_t q = nullptr;
    if (hipMalloc(&w, sizeof(double) * cg_ws_doubles(m->n)) != hipSuccess ||
        hipMalloc(&r, sizeof(double) * kRedSize) != hipSuccess ||
        hipMalloc(&c, sizeof(McgState)) != hipSuccess ||
        hipMalloc(&fl, (size_t)std::max<int64_t>(m->gs.ngroups, 1)) != hipSuccess ||
        hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      for (void *v : {(void *)w, (void *)r, (void *)c, (void *)fl}) (void)hipFree(v);
      if (q) (void)hipStreamDestroy(q);
      set_error("nkh_batch_create: out of device memory");
      return fail(NKH_ERR_NOMEM);
    }
    B->mesh.push_back(meshes[k]);
    B->ws.push_back(w);
    B->red.push_back(r);
    B->st.push_back(c);
    B->flag.push_back(fl);
    B->nflag.push_back(m->gs.ngroups);
    B->streams.push_back(q);
  }
  for (int k = 0; k <= nb; ++k) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      set_error("nkh_batch_create: event creation failed");
      return fail(NKH_ERR_HIP);
    }
    B->ev.push_back(e);
  }
  if (hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&B->ev_sync, hipEventDisableTiming) != hipSuccess) {
    set_error("nkh_batch_create: stream creation failed");
    return fail(NKH_ERR_HIP);
  }
  g_batches.push_back(B);
  *id = (int)g_batches.size() - 1;
  return NKH_OK;
}

int batch_meshes(int id, std::vector<int> *meshes) {
  NK_CHECK(id >= 0 && id < (int)g_batches.size() && g_batches[(size_t)id], NKH_ERR_ARG,
           "batch: bad handle %d", id);
  *meshes = g_batches[(size_t)id]->mesh;
  return NKH_OK;
}

int batch_destroy(int id) {
  NK_CHECK(id >= 0 && id < (int)g_batches.size() && g_batches[(size_t)id], NKH_ERR_ARG,
           "nkh_batch_destroy: bad handle %d", id);
  (void)hipDeviceSynchronize();
  g_batches[(size_t)id]->release();
  delete g_batches[(size_t)id];
  g_batches[(size_t)id] = nullptr;
  return NKH_OK;
}

namespace {

template <class Run>
int batch_drive(CgBatch &B, const double *const *x, const double *const *f, const Coef *h1,
                const Coef *h2, const double *const *mask, const double *const *mult,
                const double *tin, int maxit, const nkh_params &prm, nkh_result *res,
                int *launches_out) {
  hipStream_t s = cur_stream();
  std::vector<Run> runs((size_t)B.nb);
  std::vector<int> active;
  for (int k = 0; k < B.nb; ++k) {
    Mesh *m = get_mesh(B.mesh[(size_t)k]);
    NK_CHECK(m && m->n > 0, NKH_ERR_STATE, "batch: mesh of solve %d is gone", k);
    NK_CHECK(m->gs.ngroups == B.nflag[(size_t)k], NKH_ERR_STATE,
             "batch: the gather-scatter map of solve %d changed", k);
    // set-up of solve k on s (it uses the mesh's scratch reductions)
    NK_TRY(runs[(size_t)k].begin(*m, B.ws[(size_t)k], B.red[(size_t)k], B.st[(size_t)k],
                                 const_cast<double *>(x[k]), f[k], h1[k], h2[k], mask[k],
                                 mult[k], tin[k], maxit, prm, false, s));
    if (runs[(size_t)k].trivial) continue;
    if (runs[(size_t)k].fused && m->gs.ngroups > 0) {
      NK_HIP(hipMemcpyAsync(B.flag[(size_t)k], m->gs.flag, (size_t)m->gs.ngroups,
                            hipMemcpyDeviceToDevice, s));
      runs[(size_t)k].flag = B.flag[(size_t)k];
    }
    active.push_back(k);
  }
  // the first iteration (no pending update) eagerly, then the graph; all of
  // it on the batch's own stream c, after the set-up on s
  hipStream_t c = B.cap;
  NK_HIP(hipEventRecord(B.ev_sync, s));
  NK_HIP(hipStreamWaitEvent(c, B.ev_sync, 0));
  auto fork_join = [&](auto &&body) -> int {
    NK_HIP(hipEventRecord(B.ev[0], c));
    for (int k : active) {
      hipStream_t q = B.streams[(size_t)k];
      NK_HIP(hipStreamWaitEvent(q, B.ev[0], 0));
      NK_TRY(body(k, q));
      NK_HIP(hipEventRecord(B.ev[(size_t)k + 1], q));
    }
    for (int k : active) NK_HIP(hipStreamWaitEvent(c, B.ev[(size_t)k + 1], 0));
    return NKH_OK;
  };
  int launches = 0;
  if (!active.empty()) {
    NK_TRY(fork_join([&](int k, hipStream_t q) { return runs[(size_t)k].enqueue(q); }));
    launches = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the graph: kGraphIters iterations of every solve, each solve on its own
    // stream (solves do not wait for each other between iterations)
    NK_HIP(hipStreamBeginCapture(c, hipStreamCaptureModeThreadLocal));
    const int rc_cap = fork_join([&](int k, hipStream_t q) {
      for (int it = 0; it < kGraphIters; ++it) NK_TRY(runs[(size_t)k].enqueue(q));
      return NKH_OK;
    });
    const hipError_t end = hipStreamEndCapture(c, &graph);
    NK_TRY(rc_cap);
    NK_HIP(end);
    NK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    int niter_max = 0;
    for (int k : active) niter_max = std::max(niter_max, runs[(size_t)k].niter);
    std::vector<CgState> hs((size_t)B.nb);
    int rc = NKH_OK;
    while (true) {
      if (hipGraphLaunch(exec, c) != hipSuccess) {
        set_error("batch: graph launch failed");
        rc = NKH_ERR_HIP;
        break;
      }
      launches += kGraphIters;
      {
        for (int k : active)
          (void)hipMemcpyAsync(&hs[(size_t)k], B.st[(size_t)k], sizeof(CgState),
                               hipMemcpyDeviceToHost, c);
        if (hipStreamSynchronize(c) != hipSuccess) {
          set_error("batch: state read-back failed");
          rc = NKH_ERR_HIP;
          break;
        }
        bool all = true;
        for (int k : active) all = all && hs[(size_t)k].done;
        if (all) break;
        if (launches > niter_max + kGraphIters) {
          set_error("batch: state machines did not terminate");
          rc = NKH_ERR_STATE;
          break;
        }
      }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    NK_TRY(rc);
    NK_HIP(hipEventRecord(B.ev_sync, c));
    NK_HIP(hipStreamWaitEvent(s, B.ev_sync, 0));
    for (int k : active) runs[(size_t)k].result(hs[(size_t)k], &res[k]);
  }
  for (int k = 0; k < B.nb; ++k)
    if (runs[(size_t)k].trivial) runs[(size_t)k].result(CgState{}, &res[k]);
  if (launches_out) *launches_out = launches;
  return NKH_OK;
}

}  // namespace

int run_cggo_batch(int id, const double *const *x, const double *const *f, const Coef *h1,
                   const Coef *h2, const double *const *mask, const double *const *mult,
                   const double *tin, int maxit, const nkh_params &prm, nkh_result *res,
                   int *launches_out) {
  NK_CHECK(id >= 0 && id < (int)g_batches.size() && g_batches[(size_t)id], NKH_ERR_ARG,
           "batch: bad handle %d", id);
  CgBatch &B = *g_batches[(size_t)id];
  if (prm.precision == 2)
    return batch_drive<McgRun>(B, x, f, h1, h2, mask, mult, tin, maxit, prm, res, launches_out);
  return batch_drive<CgRun>(B, x, f, h1, h2, mask, mult, tin, maxit, prm, res, launches_out);
}

}  // namespace nekamd
