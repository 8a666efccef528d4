// Instantiations of the fused on-chip two-pass transform and its helpers.
#include "kfa_fused_impl.h"
#include "kfa_kernels.h"

namespace kfa {

hipError_t launch_fused_f32(uint32_t N, const FusedArgs& f, uint32_t groups, hipStream_t s) {
    if (N == 256)
        return launch_fused_t<float, 256, 32>(f, groups, s);
    return hipErrorNotSupported;
}

hipError_t launch_fused_f64(uint32_t N, const FusedArgs& f, uint32_t groups, hipStream_t s) {
    if (N == 256)
        return launch_fused_t<double, 256, 16>(f, groups, s);
    return hipErrorNotSupported;
}

hipError_t launch_xcc_probe(uint32_t* out, uint32_t groups, hipStream_t s) {
    hipLaunchKernelGGL(k_xcc_probe, dim3(groups), dim3(64), 0, s, out);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

} // namespace kfa
