// Single-precision instantiations of the power-of-two fast passes.
#include "kfa_fast_impl.h"
#include "kfa_kernels.h"

namespace kfa {
hipError_t launch_fast_f32(uint32_t L, int mode, const FastArgs& a, hipStream_t s) {
    return launch_fast_any<float>(L, mode, a, s);
}
uint32_t fast_tout_ipw_f32(uint32_t L) { return fast_tout_ipw_any<float>(L); }
} // namespace kfa
