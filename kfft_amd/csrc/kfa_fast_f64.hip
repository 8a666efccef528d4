// Double-precision instantiations of the power-of-two fast passes.
#include "kfa_fast_impl.h"
#include "kfa_kernels.h"

namespace kfa {
hipError_t launch_fast_f64(uint32_t L, int mode, const FastArgs& a, hipStream_t s) {
    return launch_fast_any<double>(L, mode, a, s);
}
uint32_t fast_tout_ipw_f64(uint32_t L) { return fast_tout_ipw_any<double>(L); }
} // namespace kfa
