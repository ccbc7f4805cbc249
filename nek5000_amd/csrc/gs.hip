// Gather-scatter (dssum / dsop) on one GPU.
//
// Reference semantics: dssum.f:33-98 calls fgslib_gs_op(gsh,u,1,1,0): every
// local node with nonzero glo_num receives the sum over all local nodes with
// the same id; glo_num == 0 nodes are untouched (setvert3d gives element
// interiors id 0, navier8.f:2337-2347).
//
// Layout: at setup the local nodes are grouped by nonzero id on the host
// (off/idx, int32; members in increasing local index; groups in the order of
// their first member, so that neighbouring groups lie in the same few
// elements).  Runs of whole groups of at most 64 members form packets, one
// wavefront each: lane l takes member l of the packet (one coalesced index
// load, one gather), and a group's sum is formed across its lanes by
// cross-lane reads in member order.  Every group is owned by one wavefront
// and summed in member order, so results are deterministic, independent of
// the launch shape and need no atomics.  Groups of more than 64 members
// (none on conforming hexahedral meshes of valence <= 64) are kept after the
// packets and reduced one thread per group.
#include <algorithm>
#include <climits>
#include <numeric>
#include <type_traits>
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
  // groups of more than 64 members after all others; then the packets
  int64_t nbig = 0;
  for (size_t g = 0; g + 1 < off.size(); ++g) nbig += (off[g + 1] - off[g] > 64);
  if (nbig > 0) {
    std::vector<int32_t> off2, idx2;
    off2.reserve(off.size());
    idx2.reserve(idx.size());
    for (int pass = 0; pass < 2; ++pass)
      for (size_t g = 0; g + 1 < off.size(); ++g) {
        const int32_t a = off[g], c = off[g + 1];
        if ((c - a > 64) != (pass == 1)) continue;
        off2.push_back((int32_t)idx2.size());
        idx2.insert(idx2.end(), idx.begin() + a, idx.begin() + c);
      }
    off2.push_back((int32_t)idx2.size());
    off.swap(off2);
    idx.swap(idx2);
  }
  const int64_t gpk = (int64_t)off.size() - 1 - nbig;
  std::vector<int32_t> pk_m, pk_g;
  std::vector<uint64_t> pk_head;
  for (int64_t g = 0; g < gpk;) {
    const int32_t q0 = off[(size_t)g];
    uint64_t hm = 0;
    pk_m.push_back(q0);
    pk_g.push_back((int32_t)g);
    while (g < gpk && off[(size_t)g + 1] - q0 <= 64) {
      hm |= 1ull << (off[(size_t)g] - q0);
      ++g;
    }
    pk_head.push_back(hm);
  }
  pk_m.push_back(off[(size_t)gpk]);
  pk_g.push_back((int32_t)gpk);
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
  gs.npk = (int64_t)pk_head.size();
  gs.gpk = gpk;
  auto up = [](auto **dst, const auto &v) -> int {
    using T = typename std::decay_t<decltype(v)>::value_type;
    NK_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty())
      NK_HIP(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return NKH_OK;
  };
  NK_TRY(up(&gs.off, off));
  if (!idx.empty()) NK_TRY(up(&gs.idx, idx));
  NK_TRY(up(&gs.pk_m, pk_m));
  NK_TRY(up(&gs.pk_g, pk_g));
  NK_TRY(up(&gs.pk_head, pk_head));
  return NKH_OK;
}

void gs_free(GsMap &gs) {
  if (gs.flag) (void)hipFree(gs.flag);
  if (gs.off) (void)hipFree(gs.off);
  if (gs.idx) (void)hipFree(gs.idx);
  if (gs.pk_m) (void)hipFree(gs.pk_m);
  if (gs.pk_g) (void)hipFree(gs.pk_g);
  if (gs.pk_head) (void)hipFree(gs.pk_head);
  gs = GsMap();
}

template <int OP>
__device__ __forceinline__ double gs_combine(double a, double b) {
  if constexpr (OP == 0 || OP == 4) return a + b;
  else if constexpr (OP == 1) return a * b;
  else if constexpr (OP == 2) return b < a ? b : a;
  else return b > a ? b : a;
}

// ---- packets: one wavefront per packet, lane l on member l --------------
// Packet p of a launch of 4-wavefront blocks; the blocks resident on one XCD
// (workgroups are dealt to the 8 XCDs round-robin) take neighbouring
// packets and share their L2 lines.  Wavefront-uniform (held in a scalar
// register, so the packet's descriptors are scalar loads).
__device__ __forceinline__ int pk_index() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int lb = (b % 8) * (nb / 8) + b / 8;
  return __builtin_amdgcn_readfirstlane(lb * 4 + (int)(threadIdx.x >> 6));
}

// This lane's member of packet p: node index, first lane of its group,
// group length (0 past the packet's members) and group rank in the packet.
struct PkLane {
  int node, head, len, local;
  bool act;
};
__device__ __forceinline__ PkLane pk_lane(const int32_t *__restrict__ pk_m,
                                          const uint64_t *__restrict__ pk_head,
                                          const int32_t *__restrict__ idx, int p) {
  const int lane = threadIdx.x & 63;
  const int q0 = pk_m[p], cnt = pk_m[p + 1] - q0;
  const uint64_t hm = pk_head[p];  // bit 0 is always set
  PkLane L;
  L.act = lane < cnt;
  L.node = L.act ? idx[q0 + lane] : 0;
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  L.head = 63 - __clzll((long long)(hm & upto));
  const uint64_t after = L.head == 63 ? 0ull : (hm & ~((2ull << L.head) - 1));
  const int end = after ? __ffsll((long long)after) - 1 : cnt;
  L.len = L.act ? end - L.head : 0;
  L.local = __popcll(hm & ((1ull << L.head) - 1));
  return L;
}

// The group's combination of v in member order (first member, then each
// next one), on every lane of the group.  The loop runs to the longest
// group of the wavefront.
template <typename T, typename F>
__device__ __forceinline__ T pk_reduce(T v, const PkLane &L, F comb) {
  T s = __shfl(v, L.head, 64);
  for (int k = 1; __ballot(k < L.len) != 0; ++k) {
    const T t = __shfl(v, L.head + k, 64);
    if (k < L.len) s = comb(s, t);
  }
  return s;
}

template <int OP>
struct GsComb {
  __device__ double operator()(double a, double b) const {
    if constexpr (OP == 0 || OP == 4) return a + b;
    else if constexpr (OP == 1) return a * b;
    else if constexpr (OP == 2) return b < a ? b : a;
    else return b > a ? b : a;
  }
};
struct AddComb {
  template <typename T>
  __device__ T operator()(T a, T b) const { return a + b; }
};

template <int OP>
__global__ void __launch_bounds__(256)
    k_gs_op_pk(double *__restrict__ u, const int32_t *__restrict__ pk_m,
               const uint64_t *__restrict__ pk_head, const int32_t *__restrict__ idx,
               int64_t npk) {
  const int p = pk_index();
  if (p >= npk) return;
  const PkLane L = pk_lane(pk_m, pk_head, idx, p);
  const double v = L.act ? u[L.node] : 0.0;
  double acc = pk_reduce(v, L, GsComb<OP>());
  // op 4 (dsavg, ic.f:1871-1897): the sum times vmult = 1/multiplicity
  if constexpr (OP == 4) acc = acc * (1.0 / (double)L.len);
  if (L.len >= 2) u[L.node] = acc;
}

// one thread per group, groups [g0, ng) (those of more than 64 members)
template <int OP>
__global__ void k_gs_op(double *__restrict__ u, const int32_t *__restrict__ off,
                        const int32_t *__restrict__ idx, int64_t g0, int64_t nseg) {
  for (int64_t s = g0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < nseg;
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

// blocks of 4 packets, a multiple of 8 (the XCD mapping of pk_index)
static int pk_blocks(int64_t npk) {
  const int64_t b = (npk + 3) / 4;
  return (int)((b + 7) / 8 * 8);
}
static_assert(sizeof(int) == 4, "packet indices are int32");

template <int OP>
static void gs_op_launch(const GsMap &gs, double *u, hipStream_t s) {
  if (gs.npk > 0)
    hipLaunchKernelGGL(k_gs_op_pk<OP>, dim3(pk_blocks(gs.npk)), dim3(256), 0, s, u, gs.pk_m,
                       gs.pk_head, gs.idx, gs.npk);
  const int64_t nbig = gs.ngroups - gs.gpk;
  if (nbig > 0)
    hipLaunchKernelGGL(k_gs_op<OP>, dim3((int)std::min<int64_t>((nbig + 255) / 256, 65536)),
                       dim3(256), 0, s, u, gs.off, gs.idx, gs.gpk, gs.ngroups);
}

int gs_op(const GsMap &gs, double *u, int op, hipStream_t s) {
  if (gs.nseg == 0) return NKH_OK;
  switch (op) {
    case 0: gs_op_launch<0>(gs, u, s); break;
    case 1: gs_op_launch<1>(gs, u, s); break;
    case 2: gs_op_launch<2>(gs, u, s); break;
    case 3: gs_op_launch<3>(gs, u, s); break;
    case 4: gs_op_launch<4>(gs, u, s); break;
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
    k_gs_assemble_pk(T *__restrict__ w, const int32_t *__restrict__ pk_m,
                     const int32_t *__restrict__ pk_g, const uint64_t *__restrict__ pk_head,
                     const int32_t *__restrict__ idx, const uint8_t *__restrict__ flag,
                     int64_t npk, const CgState *__restrict__ st) {
  if (st && st->done) return;
  const int p = pk_index();
  if (p >= npk) return;
  const PkLane L = pk_lane(pk_m, pk_head, idx, p);
  const T v = L.act ? w[L.node] : T(0);
  const uint8_t f = L.act ? flag[pk_g[p] + L.local] : uint8_t(2);
  const T acc = pk_reduce(v, L, AddComb());
  if (f == 1) w[L.node] = T(0);
  else if (f == 0) w[L.node] = acc;
}

// one thread per group, groups [g0, ngroups) in chunks
template <typename T>
__global__ void __launch_bounds__(256)
    k_gs_assemble(T *__restrict__ w, const int32_t *__restrict__ off,
                  const int32_t *__restrict__ idx, const uint8_t *__restrict__ flag,
                  int64_t g0, int64_t ngroups, int64_t chunk, const CgState *__restrict__ st) {
  if (st && st->done) return;
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t lb = (int64_t)(b % 8) * (nb / 8) + b / 8;
  const int64_t s0 = g0 + lb * chunk;
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
    k_gs_fin_pk(Z *__restrict__ z, const E *__restrict__ e, const W *__restrict__ w,
                const int32_t *__restrict__ pk_m, const uint64_t *__restrict__ pk_head,
                const int32_t *__restrict__ idx, int64_t npk, const double *__restrict__ lut) {
  const int p = pk_index();
  if (p >= npk) return;
  const PkLane L = pk_lane(pk_m, pk_head, idx, p);
  const E ev = L.act ? e[L.node] : E(0);
  const double wv = L.act ? wdec(w, L.node, lut) : 0.0;
  const double zv = L.act ? double(z[L.node]) : 0.0;
  const double S = double(pk_reduce(ev, L, AddComb()));
  double acc = pk_reduce(S * wv + zv, L, AddComb());
  acc = acc * (1.0 / (double)L.len);
  if (L.act) z[L.node] = Z(acc);
}

template <typename E, typename Z, typename W>
__global__ void __launch_bounds__(256)
    k_gs_fin(Z *__restrict__ z, const E *__restrict__ e, const W *__restrict__ w,
             const int32_t *__restrict__ off, const int32_t *__restrict__ idx, int64_t g0,
             int64_t ngroups, int64_t chunk, const double *__restrict__ lut) {
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t lb = (int64_t)(b % 8) * (nb / 8) + b / 8;
  const int64_t s0 = g0 + lb * chunk;
  const int64_t s1 = (s0 + chunk < ngroups) ? s0 + chunk : ngroups;
  for (int64_t g = s0 + threadIdx.x; g < s1; g += blockDim.x) {
    const int32_t a = off[g], c = off[g + 1];
    E Se = e[idx[a]];
    for (int32_t q = a + 1; q < c; ++q) Se += e[idx[q]];
    const double S = double(Se);
    double acc = S * wdec(w, idx[a], lut) + double(z[idx[a]]);
    for (int32_t q = a + 1; q < c; ++q) {
      const int32_t l = idx[q];
      acc += S * wdec(w, l, lut) + double(z[l]);
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

// nblocks (> 0): grid of the per-group kernel of the large groups
template <typename T>
static int gs_assemble_t(const GsMap &gs, T *w, const CgState *st, hipStream_t s,
                         const uint8_t *flag, int nblocks = 0) {
  if (gs.ngroups == 0) return NKH_OK;
  const uint8_t *fl = flag ? flag : gs.flag;
  if (gs.npk > 0)
    hipLaunchKernelGGL(k_gs_assemble_pk<T>, dim3(pk_blocks(gs.npk)), dim3(256), 0, s, w,
                       gs.pk_m, gs.pk_g, gs.pk_head, gs.idx, fl, gs.npk, st);
  const int64_t nbig = gs.ngroups - gs.gpk;
  if (nbig > 0) {
    const int nb = nblocks > 0 ? nblocks : gs_blocks(nbig);
    const int64_t chunk = (nbig + nb - 1) / nb;
    hipLaunchKernelGGL(k_gs_assemble<T>, dim3(nb), dim3(256), 0, s, w, gs.off, gs.idx, fl,
                       gs.gpk, gs.ngroups, chunk, st);
  }
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
static int gs_fin_t(const GsMap &gs, Z *z, const E *e, const W *w, hipStream_t s,
                    const double *lut = nullptr) {
  if (gs.ngroups == 0) return NKH_OK;
  if (gs.npk > 0)
    hipLaunchKernelGGL((k_gs_fin_pk<E, Z, W>), dim3(pk_blocks(gs.npk)), dim3(256), 0, s, z, e,
                       w, gs.pk_m, gs.pk_head, gs.idx, gs.npk, lut);
  const int64_t nbig = gs.ngroups - gs.gpk;
  if (nbig > 0) {
    const int nb = gs_blocks(nbig);
    const int64_t chunk = (nbig + nb - 1) / nb;
    hipLaunchKernelGGL((k_gs_fin<E, Z, W>), dim3(nb), dim3(256), 0, s, z, e, w, gs.off,
                       gs.idx, gs.gpk, gs.ngroups, chunk, lut);
  }
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

int gs_fin(const GsMap &gs, double *z, const double *e, const uint8_t *w, const double *lut,
           hipStream_t s) {
  return gs_fin_t(gs, z, e, w, s, lut);
}

int gs_fin(const GsMap &gs, float *z, const float *e, const uint8_t *w, const double *lut,
           hipStream_t s) {
  return gs_fin_t(gs, z, e, w, s, lut);
}

}  // namespace nekamd
