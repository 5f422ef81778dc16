// backward.hip -- placeholder until the MFMA backward lands.
#include "sep_internal.hpp"

namespace sep {

int eig_backward_f64(sep_handle*, int64_t, const double*, const double*, int64_t, const double*,
                     const double*, double*, double*)
{
    set_error("sep_eig_backward_f64_dev: not built yet");
    return SEP_ERR_UNSUPPORTED;
}
int eig_backward_batched_f64(sep_handle*, int64_t, const int64_t*, const int64_t*,
                             const int64_t*, const double*, const double*, const double*,
                             const double*, double*, double*)
{
    set_error("sep_eig_backward_batched_f64_dev: not built yet");
    return SEP_ERR_UNSUPPORTED;
}

}  // namespace sep
