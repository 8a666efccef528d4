// Single-precision instantiations of the LDS FFT pass.
#include "kfa_fft_impl.h"
#include "kfa_kernels.h"

namespace kfa {
hipError_t launch_fft_f32(int maxe, const KfaFftArgs& a, uint32_t lds_bytes, hipStream_t s) {
    return maxe <= 16 ? launch_fft_impl<float, 16>(a, lds_bytes, s)
                      : launch_fft_impl<float, 32>(a, lds_bytes, s);
}
} // namespace kfa
