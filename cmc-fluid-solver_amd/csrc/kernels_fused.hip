// Fused line sweep (placeholder until the wavefront-per-line kernels land).
#include "cmcx_internal.h"

namespace cmcx {

template <class T>
bool fused_supported(const Geom&) {
    return false;
}

template <class T>
void launch_fused_sweep(const Geom&, int, const SweepParams<T>&, const uint8_t*, const uint8_t*,
                        const uint8_t*, const T* const*, Layer4<T>, Layer4<T>, Layer4<T>,
                        Layer4<T>, int, const int32_t*, int64_t, hipStream_t) {
    throw std::string("fused sweep not available");
}

void build_active_lines(const uint8_t*, const Geom&, int, int32_t** d_lines, int64_t* nlines,
                        hipStream_t) {
    *d_lines = nullptr;
    *nlines = 0;
}

template bool fused_supported<float>(const Geom&);
template bool fused_supported<double>(const Geom&);
template void launch_fused_sweep<float>(const Geom&, int, const SweepParams<float>&, const uint8_t*,
                                        const uint8_t*, const uint8_t*, const float* const*,
                                        Layer4<float>, Layer4<float>, Layer4<float>, Layer4<float>,
                                        int, const int32_t*, int64_t, hipStream_t);
template void launch_fused_sweep<double>(const Geom&, int, const SweepParams<double>&,
                                         const uint8_t*, const uint8_t*, const uint8_t*,
                                         const double* const*, Layer4<double>, Layer4<double>,
                                         Layer4<double>, Layer4<double>, int, const int32_t*,
                                         int64_t, hipStream_t);

}  // namespace cmcx
