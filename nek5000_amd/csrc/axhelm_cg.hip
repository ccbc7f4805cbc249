// axhelm fused into the CG iteration: instantiations for lx1 = 2..16.
#include "element_kernels.h"

namespace nekamd {

int launch_axhelm_cg(const Mesh &m, double *w, double *p, const double *r, const double *d,
                     Coef h1, Coef h2, const double *mult, const CgState *st,
                     double *rho_part, hipStream_t s) {
#define CALL(NP) axhelm_dispatch<NP, 1>(m, w, p, h1, h2, p, r, d, mult, st, rho_part, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

int launch_axhelm_flex(const Mesh &m, double *w, double *p, const double *z, Coef h1, Coef h2,
                       const double *betap, double *rho_part, hipStream_t s, double *x,
                       const double *alphap) {
#define CALL(NP)                                                                            \
  axhelm_dispatch<NP, 2>(m, w, p, h1, h2, p, z, nullptr, nullptr, nullptr, rho_part, s, betap, \
                         x, alphap)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

}  // namespace nekamd
