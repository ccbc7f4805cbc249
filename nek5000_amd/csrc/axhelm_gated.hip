// The plain axhelm gated by a device state: a captured iteration replays it
// every pass, and it does work only on the passes whose state is open
// (the FP64 true-residual refresh of the mixed-precision Jacobi PCG).
#include "element_kernels.h"

namespace nekamd {

int launch_axhelm_gated(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                        const CgState *gate, hipStream_t s) {
#define CALL(NP) \
  axhelm_dispatch<NP, 3>(m, au, u, h1, h2, nullptr, nullptr, nullptr, nullptr, gate, nullptr, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

}  // namespace nekamd
