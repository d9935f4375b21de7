// FP32 instantiation of the fused time-step kernels (sles_rhs.h).
#include "sles_rhs.h"

namespace sles {

int step_f32(const sles_plan_s* p, int scheme, const sles_step_params* prm, void* uh, void* uh0, void* uh1,
             void* aux, long ps, int ssplit, void* u_out, void* curl_out, double* scalars, void* work,
             hipStream_t s, void* stages, const DriveArgs* drive, const Batch* batch) {
#define SLES_CASE(NN) \
  case NN:            \
    return step_t<float, NN>(p, scheme, prm, uh, uh0, uh1, aux, ps, ssplit, u_out, curl_out, scalars, work, s, stages, \
                         drive, batch);
  switch (p->n) {
    SLES_FOR_EACH_N(SLES_CASE)
    default:
      return fail("sles_step: unsupported grid size");
  }
#undef SLES_CASE
}

}  // namespace sles
