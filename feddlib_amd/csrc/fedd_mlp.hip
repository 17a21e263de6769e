// Learned tangent surrogate (placeholder until the network kernels land).
#include "fedd_internal.h"

namespace fedd {

int launch_mlp_tangent(Ctx* c, int nq) {
  (void)nq;
  return fail(c, FEDD_ERR_STATE, "no surrogate network loaded");
}

}  // namespace fedd
