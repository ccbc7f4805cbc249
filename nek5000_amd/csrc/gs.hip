// Gather-scatter (dssum / dsop) on one GPU.
//
// Reference semantics: dssum.f:33-98 calls fgslib_gs_op(gsh,u,1,1,0): every
// local node with nonzero glo_num receives the sum over all local nodes with
// the same id; glo_num == 0 nodes are untouched (setvert3d gives element
// interiors id 0, navier8.f:2337-2347).
//
// Layout: at setup the nonzero ids are sorted (id, local index) on the host;
// groups of >= 2 members become segments (off/idx, int32, ordered by id so
// that the segments of one face or edge are adjacent), unpartnered ids become
// `single`.  One thread reduces one segment in index order, so results are
// deterministic and independent of the launch shape; no atomics are needed
// because every segment is owned by exactly one thread.
#include <algorithm>
#include <climits>
#include <numeric>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

int gs_build(GsMap &gs, const int64_t *glo, int64_t n, int np, int64_t nel) {
  NK_CHECK(n < ((int64_t)1 << 31), NKH_ERR_ARG, "gs: %lld local nodes exceed int32 indexing",
           (long long)n);
  gs_free(gs);
  // Group the nonzero ids (members in increasing local index) and list the
  // groups in the order of their first member, so that the groups one block
  // of threads visits lie in the same few elements.  Ids from setvert3d are
  // dense in [1, ngv]: a counting sort does it in O(n); arbitrary ids fall
  // back to a comparison sort.
  std::vector<int32_t> off, idx;
  int maxm = 0;
  int64_t nsingle = 0;
  int64_t gmax = 0, gmin = INT64_MAX, nnz = 0;
  for (int64_t i = 0; i < n; ++i)
    if (glo[i] != 0) {
      gmax = std::max(gmax, glo[i]);
      gmin = std::min(gmin, glo[i]);
      ++nnz;
    }
  off.reserve((size_t)nnz / 2 + 2);
  idx.reserve((size_t)nnz);
  auto emit = [&](const int32_t *mem, int64_t cnt) {
    off.push_back((int32_t)idx.size());
    idx.insert(idx.end(), mem, mem + cnt);
    maxm = std::max(maxm, (int)cnt);
    if (cnt == 1) ++nsingle;
  };
  if (nnz > 0 && gmin > 0 && gmax <= 4 * n + 16) {
    std::vector<int32_t> cnt((size_t)gmax + 2, 0);
    for (int64_t i = 0; i < n; ++i)
      if (glo[i] != 0) ++cnt[(size_t)glo[i] + 1];
    for (size_t k = 1; k < cnt.size(); ++k) cnt[k] += cnt[k - 1];
    std::vector<int32_t> perm((size_t)nnz);
    {
      std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        if (glo[i] != 0) perm[(size_t)pos[(size_t)glo[i]]++] = (int32_t)i;
    }
    for (int64_t l = 0; l < n; ++l) {
      const int64_t g = glo[l];
      if (g != 0 && perm[(size_t)cnt[(size_t)g]] == (int32_t)l)
        emit(perm.data() + cnt[(size_t)g], cnt[(size_t)g + 1] - cnt[(size_t)g]);
    }
  } else if (nnz > 0) {
    std::vector<std::pair<int64_t, int32_t>> p;
    p.reserve((size_t)nnz);
    for (int64_t i = 0; i < n; ++i)
      if (glo[i] != 0) p.emplace_back(glo[i], (int32_t)i);
    std::sort(p.begin(), p.end());
    std::vector<std::pair<int32_t, size_t>> starts;  // (first member, position in p)
    for (size_t a = 0; a < p.size();) {
      size_t b = a;
      while (b < p.size() && p[b].first == p[a].first) ++b;
      starts.emplace_back(p[a].second, a);
      a = b;
    }
    std::sort(starts.begin(), starts.end());
    std::vector<int32_t> mem;
    for (auto &st : starts) {
      mem.clear();
      for (size_t b = st.second; b < p.size() && p[b].first == p[st.second].first; ++b)
        mem.push_back(p[b].second);
      emit(mem.data(), (int64_t)mem.size());
    }
  }
  off.push_back((int32_t)idx.size());
  // structural check used by the fused CG kernels: ids are zero exactly at
  // element-interior nodes (then those nodes can be finished inside axhelm)
  bool iz = (n == (int64_t)np * np * np * nel);
  if (iz) {
    const int np3 = np * np * np;
    for (int64_t l = 0; l < n && iz; ++l) {
      const int q = (int)(l % np3);
      const int i = q % np, j = (q / np) % np, k = q / (np * np);
      const bool interior = i > 0 && i < np - 1 && j > 0 && j < np - 1 && k > 0 && k < np - 1;
      if (interior != (glo[l] == 0)) iz = false;
    }
  }
  gs.n = n;
  gs.ngroups = (int64_t)off.size() - 1;
  gs.nmem = (int64_t)idx.size();
  gs.nseg = gs.ngroups - nsingle;
  gs.nsingle = nsingle;
  gs.interior_zero = iz;
  gs.max_mult = maxm;
  NK_HIP(hipMalloc(&gs.off, sizeof(int32_t) * off.size()));
  NK_HIP(hipMemcpy(gs.off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
  if (!idx.empty()) {
    NK_HIP(hipMalloc(&gs.idx, sizeof(int32_t) * idx.size()));
    NK_HIP(hipMemcpy(gs.idx, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice));
  }
  return NKH_OK;
}

void gs_free(GsMap &gs) {
  if (gs.flag) (void)hipFree(gs.flag);
  if (gs.off) (void)hipFree(gs.off);
  if (gs.idx) (void)hipFree(gs.idx);
  gs = GsMap();
}

template <int OP>
__device__ __forceinline__ double gs_combine(double a, double b) {
  if constexpr (OP == 0 || OP == 4) return a + b;
  else if constexpr (OP == 1) return a * b;
  else if constexpr (OP == 2) return b < a ? b : a;
  else return b > a ? b : a;
}

template <int OP>
__global__ void k_gs_op(double *__restrict__ u, const int32_t *__restrict__ off,
                        const int32_t *__restrict__ idx, int64_t nseg) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = off[s], b = off[s + 1];
    if (b - a < 2) continue;
    double acc = u[idx[a]];
    for (int32_t c = a + 1; c < b; ++c) acc = gs_combine<OP>(acc, u[idx[c]]);
    // op 4 (dsavg, ic.f:1871-1897): the sum times vmult = 1/multiplicity
    if constexpr (OP == 4) acc = acc * (1.0 / (double)(b - a));
    for (int32_t c = a; c < b; ++c) u[idx[c]] = acc;
  }
}

int gs_op(const GsMap &gs, double *u, int op, hipStream_t s) {
  if (gs.nseg == 0) return NKH_OK;
  const int64_t ng = gs.ngroups;
  int grid = (int)((ng + 255) / 256);
  if (grid > 65536) grid = 65536;
  switch (op) {
    case 0: hipLaunchKernelGGL(k_gs_op<0>, dim3(grid), dim3(256), 0, s, u, gs.off, gs.idx, ng); break;
    case 1: hipLaunchKernelGGL(k_gs_op<1>, dim3(grid), dim3(256), 0, s, u, gs.off, gs.idx, ng); break;
    case 2: hipLaunchKernelGGL(k_gs_op<2>, dim3(grid), dim3(256), 0, s, u, gs.off, gs.idx, ng); break;
    case 3: hipLaunchKernelGGL(k_gs_op<3>, dim3(grid), dim3(256), 0, s, u, gs.off, gs.idx, ng); break;
    case 4: hipLaunchKernelGGL(k_gs_op<4>, dim3(grid), dim3(256), 0, s, u, gs.off, gs.idx, ng); break;
    default: set_error("gs_op: unknown op %d", op); return NKH_ERR_ARG;
  }
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

// CG assembly of w (hmholtz.f:797-798, dssum then mask), with the group
// flags of gs_prepare: 0 sum the members, 1 masked group (write zeros),
// 2 nothing to do (one member, mask 1).  Each block takes a contiguous chunk
// of groups; chunks are handed out so that the blocks resident on one XCD
// (workgroups are dealt to the 8 XCDs round-robin) walk neighbouring chunks
// and share their L2 lines.
template <typename T>
__global__ void __launch_bounds__(256)
    k_gs_assemble(T *__restrict__ w, const int32_t *__restrict__ off,
                  const int32_t *__restrict__ idx, const uint8_t *__restrict__ flag,
                  int64_t ngroups, int64_t chunk, const CgState *__restrict__ st) {
  if (st && st->done) return;
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t lb = (int64_t)(b % 8) * (nb / 8) + b / 8;
  const int64_t s0 = lb * chunk;
  const int64_t s1 = (s0 + chunk < ngroups) ? s0 + chunk : ngroups;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    const uint8_t f = flag[s];
    if (f == 2) continue;
    const int32_t a = off[s], e = off[s + 1];
    if (f == 1) {
      for (int32_t q = a; q < e; ++q) w[idx[q]] = T(0);
    } else {
      T acc = w[idx[a]];
      for (int32_t q = a + 1; q < e; ++q) acc += w[idx[q]];
      for (int32_t q = a; q < e; ++q) w[idx[q]] = acc;
    }
  }
}

// Last step of the multigrid V-cycle at the fine level, per group of shared
// nodes: S = dssum of the Schwarz sum e (hsmg.f h1mg_schwarz's dssum), then
// the weighted value plus the coarse correction already in z,
// S*w_i + z_i, and dsavg of that (ic.f:1871-1897: dssum, times
// 1/multiplicity).  Every id group, single members included.
template <typename E, typename Z, typename W>
__global__ void __launch_bounds__(256)
    k_gs_fin(Z *__restrict__ z, const E *__restrict__ e, const W *__restrict__ w,
             const int32_t *__restrict__ off, const int32_t *__restrict__ idx, int64_t ngroups,
             int64_t chunk) {
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t lb = (int64_t)(b % 8) * (nb / 8) + b / 8;
  const int64_t s0 = lb * chunk;
  const int64_t s1 = (s0 + chunk < ngroups) ? s0 + chunk : ngroups;
  for (int64_t g = s0 + threadIdx.x; g < s1; g += blockDim.x) {
    const int32_t a = off[g], c = off[g + 1];
    E Se = e[idx[a]];
    for (int32_t q = a + 1; q < c; ++q) Se += e[idx[q]];
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

// Group flags from mask and mult, plus the check that makes the fused CG
// exact: in every group mask is 0 or 1 and equal at all members, and the
// members' mult sum to 1 (then sum_l mult*w*p over the assembled, masked w
// equals sum_l w_local*p for a masked, continuous p).  part[b] = 0 when the
// block's groups pass.
__global__ void __launch_bounds__(256)
    k_gs_prepare(const double *__restrict__ mask, const double *__restrict__ mult,
                 const int32_t *__restrict__ off, const int32_t *__restrict__ idx,
                 int64_t ngroups, uint8_t *__restrict__ flag, double *__restrict__ part) {
  __shared__ double sbuf[4];
  double bad = 0.0;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < ngroups;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = off[s], e = off[s + 1];
    const double m0 = mask[idx[a]];
    double sm = 0.0;
    for (int32_t q = a; q < e; ++q) {
      const int32_t l = idx[q];
      if (mask[l] != m0) bad = 1.0;
      sm += mult[l];
    }
    if (!(m0 == 0.0 || m0 == 1.0) || sm != 1.0) bad = 1.0;
    flag[s] = (m0 == 0.0) ? 1 : ((e - a >= 2) ? 0 : 2);
  }
  for (int o = 32; o > 0; o >>= 1) bad = fmax(bad, __shfl_down(bad, o, 64));
  if ((threadIdx.x & 63) == 0) sbuf[threadIdx.x >> 6] = bad;
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = fmax(fmax(sbuf[0], sbuf[1]), fmax(sbuf[2], sbuf[3]));
}

int gs_prepare(GsMap &gs, const double *mask, const double *mult, double *red, bool *ok,
               hipStream_t s) {
  *ok = false;
  if (gs.ngroups == 0) return NKH_OK;
  if (!gs.flag) NK_HIP(hipMalloc(&gs.flag, (size_t)gs.ngroups));
  hipLaunchKernelGGL(k_gs_prepare, dim3(kRedBlocks), dim3(256), 0, s, mask, mult, gs.off,
                     gs.idx, gs.ngroups, gs.flag, red);
  NK_HIP(hipGetLastError());
  std::vector<double> h(kRedBlocks);
  NK_HIP(hipMemcpyAsync(h.data(), red, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  double bad = 0.0;
  for (double v : h) bad = std::max(bad, v);
  *ok = (bad == 0.0);
  return NKH_OK;
}

// Blocks of the chunked group kernels: about 512 groups per block, 64 to
// kGsBlocks, a multiple of 8 (their XCD mapping); small maps would otherwise
// pay mostly for dispatching kGsBlocks blocks.  Results do not depend on it.
static int gs_blocks(int64_t ngroups) {
  int64_t b = (ngroups + 511) / 512;
  b = (b + 7) / 8 * 8;
  return (int)(b < 64 ? 64 : (b > kGsBlocks ? kGsBlocks : b));
}

template <typename T>
static int gs_assemble_t(const GsMap &gs, T *w, const CgState *st, hipStream_t s,
                         const uint8_t *flag, int nblocks = 0) {
  if (gs.ngroups == 0) return NKH_OK;
  const int nb = nblocks > 0 ? nblocks : gs_blocks(gs.ngroups);
  const int64_t chunk = (gs.ngroups + nb - 1) / nb;
  hipLaunchKernelGGL(k_gs_assemble<T>, dim3(nb), dim3(256), 0, s, w, gs.off, gs.idx,
                     flag ? flag : gs.flag, gs.ngroups, chunk, st);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int gs_assemble(const GsMap &gs, double *w, const CgState *st, hipStream_t s,
                const uint8_t *flag, int nblocks) {
  return gs_assemble_t(gs, w, st, s, flag, nblocks);
}

int gs_assemble(const GsMap &gs, float *w, const CgState *st, hipStream_t s,
                const uint8_t *flag) {
  return gs_assemble_t(gs, w, st, s, flag);
}

template <typename E, typename Z, typename W>
static int gs_fin_t(const GsMap &gs, Z *z, const E *e, const W *w, hipStream_t s) {
  if (gs.ngroups == 0) return NKH_OK;
  const int nb = gs_blocks(gs.ngroups);
  const int64_t chunk = (gs.ngroups + nb - 1) / nb;
  hipLaunchKernelGGL((k_gs_fin<E, Z, W>), dim3(nb), dim3(256), 0, s, z, e, w, gs.off, gs.idx,
                     gs.ngroups, chunk);
  NK_HIP(hipGetLastError());
  return NKH_OK;
}

int gs_fin(const GsMap &gs, double *z, const double *e, const double *w, hipStream_t s) {
  return gs_fin_t(gs, z, e, w, s);
}

int gs_fin(const GsMap &gs, double *z, const float *e, const double *w, hipStream_t s) {
  return gs_fin_t(gs, z, e, w, s);
}

int gs_fin(const GsMap &gs, float *z, const float *e, const float *w, hipStream_t s) {
  return gs_fin_t(gs, z, e, w, s);
}

}  // namespace nekamd
