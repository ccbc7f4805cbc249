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

// Autotuner: on the first apply of an order on a mesh of at least 2^20
// local nodes (and not inside a graph capture), time both kernels with two
// grid sizes on the call's own data -- the plain apply is element-local, so
// every candidate writes the same au -- and keep the fastest for that order
// (also for the CG-fused kernels, whose grid stays fixed).
static int tune_axhelm(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                       hipStream_t s) {
  AxhelmChoice &ch = g_axhelm_choice[m.np];
  const AxhelmChoice cands[4] = {{0, 16384, false}, {0, 0, false}, {1, 16384, false},
                                 {1, 0, false}};
  hipEvent_t e0, e1;
  NK_HIP(hipEventCreate(&e0));
  NK_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  AxhelmChoice pick = ch;
  for (const AxhelmChoice &c : cands) {
    ch = c;
    NK_TRY(axhelm_apply(m, au, u, h1, h2, s));  // warm-up
    NK_HIP(hipEventRecord(e0, s));
    NK_TRY(axhelm_apply(m, au, u, h1, h2, s));
    NK_HIP(hipEventRecord(e1, s));
    NK_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    NK_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      pick = c;
    }
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
  NK_CHECK(np >= 2 && np <= kMaxNP && (variant == 0 || variant == 1) && grid >= 0,
           NKH_ERR_ARG, "axhelm tuning: lx1=%d variant=%d grid=%d", np, variant, grid);
  g_axhelm_choice[np] = AxhelmChoice{variant, grid, true};
  return NKH_OK;
}

}  // namespace nekamd
