// axhelm on its own (hmholtz.f:72): instantiations for lx1 = 2..16.
#include "element_kernels.h"

namespace nekamd {

int launch_axhelm(const Mesh &m, double *au, const double *u, Coef h1, Coef h2,
                  hipStream_t s) {
#define CALL(NP) axhelm_dispatch<NP, 0>(m, au, u, h1, h2, nullptr, nullptr, nullptr, nullptr, \
                                        nullptr, nullptr, s)
  NK_NP_SWITCH(m.np, CALL)
#undef CALL
}

}  // namespace nekamd
