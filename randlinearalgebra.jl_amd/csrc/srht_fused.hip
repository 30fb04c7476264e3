// Single-pass SRHT forward operator for the shapes the fused kernels cover.
#include "rla_internal.hpp"

namespace rla {

template <class T>
int srht_fused_forward(rla_ctx*, int64_t, int64_t, int64_t, int64_t, const int32_t*,
                       const uint64_t*, double, const T*, int64_t, int64_t, T*, int64_t, int64_t,
                       double, double, bool* handled) {
  *handled = false;
  return RLA_OK;
}

template int srht_fused_forward<float>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                       const int32_t*, const uint64_t*, double, const float*,
                                       int64_t, int64_t, float*, int64_t, int64_t, double, double,
                                       bool*);
template int srht_fused_forward<double>(rla_ctx*, int64_t, int64_t, int64_t, int64_t,
                                        const int32_t*, const uint64_t*, double, const double*,
                                        int64_t, int64_t, double*, int64_t, int64_t, double,
                                        double, bool*);

}  // namespace rla
