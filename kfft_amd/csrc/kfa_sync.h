// Intra-launch ordering between workgroups of a persistent grid that stay on
// one XCD: the XCD id, a bounded wait on a counter, and a completion signal
// issued after the workgroup's stores are acknowledged.
#pragma once
#include "kfa_device.h"

namespace kfa {

__device__ __forceinline__ uint32_t xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u; // HW_REG_XCC_ID[3:0]
}

// thread 0 only: wait until *p >= target (bounded); returns false on timeout
__device__ __forceinline__ bool wait_geq(uint32_t* p, uint32_t target, uint32_t limit,
                                         uint32_t* err) {
    for (uint32_t it = 0;; ++it) {
        const uint32_t v = __hip_atomic_fetch_add(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v >= target)
            return true;
        if (it >= limit) {
            __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

// all of this workgroup's global stores acknowledged, then one increment
__device__ __forceinline__ void signal_done(uint32_t* p) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Acquire after a completed wait: drop this CU's vector L1 lines so that loads
// see what other CUs of the XCD stored to the shared L2.  A workgroup-scope
// invalidate does not reach the L1 for data of other CUs; the agent-scope
// fence emits the device-scope invalidate (buffer_inv sc1).
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

} // namespace kfa
