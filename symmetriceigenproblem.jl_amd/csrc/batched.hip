// batched.hip -- placeholder until the persistent batched solver lands.
#include "sep_internal.hpp"

namespace sep {

int qr_batched_f64(sep_handle*, int64_t, const int64_t*, const int64_t*, const int64_t*, double*,
                   double*, double*, int64_t, int32_t*, int64_t*)
{
    set_error("sep_qr_batched_f64_dev: not built yet");
    return SEP_ERR_UNSUPPORTED;
}
int tune_batched_f64(sep_handle*, int64_t, const int64_t*, const int64_t*, const int64_t*,
                     const double*, const double*, int32_t*)
{
    set_error("sep_tune_batched_f64: not built yet");
    return SEP_ERR_UNSUPPORTED;
}

}  // namespace sep
