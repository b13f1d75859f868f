// Fast-path drivers (placeholder until the fast kernels land).
#include "pii_capi_internal.hpp"

namespace pii {

void form_factor_fast_full(Context&, const Batch&, const KGridDev&, double*) {
  throw Error(PII_ERR_UNSUPPORTED, "fast form factor not built yet");
}
void volume_matrix_fast(Context&, const Batch&, const KGridDev&, int, double*) {
  throw Error(PII_ERR_UNSUPPORTED, "fast volume matrix not built yet");
}

}  // namespace pii
