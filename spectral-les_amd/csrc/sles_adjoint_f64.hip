// FP64 instantiation of the adjoint RHS kernels (sles_adjoint.h).
#include "sles_adjoint.h"

namespace sles {

int adjoint_f64(const sles_plan_s* p, const sles_step_params* prm, const void* y, const AdjSrc& mu, void* out,
                int accumulate, int linear, double dcoeff, double* cs_grad, void* work, hipStream_t s) {
#define SLES_CASE(NN) \
  case NN:            \
    return adjoint_t<double, NN>(p, prm, y, mu, out, accumulate, linear, dcoeff, cs_grad, work, s);
  switch (p->n) {
    SLES_FOR_EACH_N(SLES_CASE)
    default:
      return fail("sles_adjoint_rhs: unsupported grid size");
  }
#undef SLES_CASE
}

}  // namespace sles
