// axhelm on its own (hmholtz.f:72): instantiations for lx1 = 2..16.
#include "element_kernels.h"

namespace nekamd {

static int axhelm_apply(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                        hipStream_t s) {
#define CALL(NP) axhelm_dispatch<NP, 0>(m, au, u, h1, h2, nullptr, nullptr, nullptr, nullptr, \
                                        nullptr, nullptr, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

static int wide_mask(int np) {
#define CALL(NP) axhelm_wide_mask<NP>()
  NK_NP_SWITCH(np, CALL)
#undef CALL
}

// Autotuner: on the first apply of an order on a mesh of at least 2^20
// local nodes (and not inside a graph capture), time every kernel variant
// the order has (element- or plane-staged slabs; the orders whose NP x NP
// slabs leave lanes idle also pack epb_wide slabs per block; NP = 8..16 also
// the matrix-core contractions, element- or plane-group-staged) with two
// grid sizes on the call's own data
// -- the plain apply is element-local, so every candidate writes au (the
// same bits, or to rounding for the matrix cores) -- and keep the fastest
// for that order; then time the variants of the CG-fused and flexible-CG-fused
// applies (their grid is fixed) on scratch vectors and keep the fastest of
// each.
static int tune_axhelm(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                       hipStream_t s) {
  AxhelmChoice &ch = g_axhelm_choice[m.np];
  hipEvent_t e0, e1;
  NK_HIP(hipEventCreate(&e0));
  NK_HIP(hipEventCreate(&e1));
  // one warm-up launch, then three timed back to back (their sum: a single
  // launch's time varies by a few per cent, as much as the variants differ)
  auto timed = [&](auto &&launch, float *ms) -> int {
    NK_TRY(launch());
    NK_HIP(hipEventRecord(e0, s));
    for (int rep = 0; rep < 3; ++rep) NK_TRY(launch());
    NK_HIP(hipEventRecord(e1, s));
    NK_HIP(hipEventSynchronize(e1));
    NK_HIP(hipEventElapsedTime(ms, e0, e1));
    return NKH_OK;
  };
  const int wm = wide_mask(m.np);
  int vars[6] = {0, 1, 0, 0, 0, 0}, nv = 2;
  if (wm & 4) vars[nv++] = 2;
  if (wm & 8) vars[nv++] = 3;
  if (wm & 16) {
    vars[nv++] = 4;
    vars[nv++] = 5;
  }
  float best = 1e30f;
  AxhelmChoice pick = ch;
  for (int c = 0; c < 2 * nv; ++c) {
    ch.variant = vars[c / 2];
    ch.grid = (c % 2 == 0) ? 16384 : 0;
    float ms = 0.f;
    NK_TRY(timed([&] { return axhelm_apply(m, au, u, h1, h2, s); }, &ms));
    if (ms < best) {
      best = ms;
      pick.variant = ch.variant;
      pick.grid = ch.grid;
    }
  }
  ch = pick;
  // the fused applies on scratch: p, r, d, w and a zeroed state (beta 0)
  const int64_t n = m.n;
  const size_t part = (size_t)kAxBlocks;
  double *buf = nullptr;
  const size_t bytes = sizeof(double) * (4 * (size_t)n + part) + sizeof(CgState);
  if (hipMalloc(&buf, bytes) == hipSuccess) {
    double *p = buf, *r = buf + n, *d = buf + 2 * n, *w = buf + 3 * n, *rho = buf + 4 * n;
    CgState *st = reinterpret_cast<CgState *>(buf + 4 * n + part);
    NK_HIP(hipMemsetAsync(buf, 0, bytes, s));
    float ms[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nv; ++c) {
      ch.variant_cg = vars[c];
      NK_TRY(timed([&] { return launch_axhelm_cg(m, w, p, r, d, h1, h2, r, st, rho, s); },
                   &ms[c]));
    }
    int bc = 0;
    for (int c = 1; c < nv; ++c)
      if (ms[c] < ms[bc]) bc = c;
    pick.variant_cg = vars[bc];
    for (int c = 0; c < nv; ++c) {
      ch.variant_flex = vars[c];
      // with the solution update the flexible CG fuses into this apply (d
      // stands in for x; alpha is 0 in the zeroed state)
      auto flex = [&] {
        return launch_axhelm_flex(m, w, p, r, h1, h2, &st->beta, rho, s, d, &st->alpha);
      };
      NK_TRY(timed(flex, &ms[c]));
    }
    bc = 0;
    for (int c = 1; c < nv; ++c)
      if (ms[c] < ms[bc]) bc = c;
    pick.variant_flex = vars[bc];
    NK_HIP(hipStreamSynchronize(s));
    (void)hipFree(buf);
  } else {
    (void)hipGetLastError();  // no room for scratch: the fused applies follow the plain one
    pick.variant_cg = pick.variant_flex = pick.variant;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pick.tuned = true;
  ch = pick;
  return NKH_OK;
}

int axhelm_autotune(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                    hipStream_t s) {
  if (m.np >= 2 && m.np <= kMaxNP && !g_axhelm_choice[m.np].tuned && m.n >= ((int64_t)1 << 20) &&
      au != u) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    NK_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) NK_TRY(tune_axhelm(m, au, u, h1, h2, s));
  }
  return NKH_OK;
}

int launch_axhelm(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                  hipStream_t s) {
  NK_TRY(axhelm_autotune(m, au, u, h1, h2, s));
  return axhelm_apply(m, au, u, h1, h2, s);
}

int axhelm_tuning(int np, int *variant, int *grid, int *tuned) {
  NK_CHECK(np >= 2 && np <= kMaxNP, NKH_ERR_ARG, "axhelm tuning: lx1=%d", np);
  const AxhelmChoice &c = g_axhelm_choice[np];
  if (variant) *variant = c.variant;
  if (grid) *grid = c.grid;
  if (tuned) *tuned = c.tuned ? 1 : 0;
  return NKH_OK;
}

int axhelm_set_tuning(int np, int variant, int grid) {
  NK_CHECK(np >= 2 && np <= kMaxNP && variant >= 0 && variant <= 5 && grid >= 0,
           NKH_ERR_ARG, "axhelm tuning: lx1=%d variant=%d grid=%d", np, variant, grid);
  g_axhelm_choice[np] = AxhelmChoice{variant, grid, variant, variant, true};
  return NKH_OK;
}

int axhelm_tuning_fused(int np, int *variant_cg, int *variant_flex) {
  NK_CHECK(np >= 2 && np <= kMaxNP, NKH_ERR_ARG, "axhelm tuning: lx1=%d", np);
  const AxhelmChoice &c = g_axhelm_choice[np];
  if (variant_cg) *variant_cg = c.variant_cg;
  if (variant_flex) *variant_flex = c.variant_flex;
  return NKH_OK;
}

}  // namespace nekamd
