/*
 * Basic types, flags and return codes.  Names and bit values are the
 * reference's public contract (incs/kfft_types.h:21-97) and must not change;
 * entries marked "GPU" are additions of this library.
 */
#pragma once
#include <inttypes.h>
#include <stdint.h>

/* incs/kfft_types.h:21-25: double by default, float under KFFT_HALF_SCALAR */
#if defined(KFFT_HALF_SCALAR)
typedef float kfft_scalar;
#else
typedef double kfft_scalar;
#endif

/* incs/kfft_types.h:28-31: interleaved complex */
typedef struct {
    kfft_scalar r;
    kfft_scalar i;
} kfft_cpx;

/* incs/kfft_types.h:34-43 */
enum kfft_info_flags {
    KFFT_INFO_TRACE = 1 << 0,
    KFFT_INFO_USE_SIMD = 1 << 1,
    KFFT_INFO_USE_ALLOCA = 1 << 2,
    KFFT_INFO_USE_SYSMATH = 1 << 3,
    KFFT_INFO_USE_OPENMP = 1 << 4,
    KFFT_INFO_RADER_ALGO = 1 << 5,
    KFFT_INFO_MEMLESS_MODE = 1 << 6,
    KFFT_INFO_HALF_SCALAR = 1 << 7,
    KFFT_INFO_GPU_HIP = 1 << 8, /* GPU: transforms run on an AMD GPU */
};

/* incs/kfft_types.h:46-57 */
enum kfft_eval_flags {
    KFFT_FLAG_NORMAL = 0,
    KFFT_FLAG_INVERSE = 1 << 0,
    KFFT_FLAG_RENEW = 1 << 1,
    KFFT_FLAG_GENERIC = 1 << 2,
    KFFT_FLAG_GENERIC_ONLY = 1 << 3,
    KFFT_FLAG_EXPAND_SCALAR = 1 << 4,
    KFFT_FLAG_DISABLE_NORM = 1 << 5,
    /* GPU: precision form of the float build (KFFT_HALF_SCALAR), see kfft_gpu.h */
    KFFT_FLAG_GPU_SPLIT_F16 = 1 << 8,
    KFFT_FLAG_GPU_SPLIT_BF16 = 1 << 9,
    KFFT_FLAG_GPU_REFINE = 1 << 10,
    /* GPU: run a 256x256 two-pass transform as one persistent launch that keeps the
     * intermediate in each XCD's L2 (experimental; library stream only) */
    KFFT_FLAG_GPU_ONCHIP = 1 << 11,
    /* GPU, ragged batches: pick each length class's plan shape with the cost
     * model set by kfft_ragged_set_cost_model, or by exhaustive timing when none
     * is set */
    KFFT_FLAG_GPU_TUNED = 1 << 12,
};

/* incs/kfft_types.h:65 */
#define KFFT_CHECK_FLAGS(X) ((X) & (~KFFT_FLAG_RENEW))

/* incs/kfft_types.h:73-78 */
typedef struct {
    uint16_t vmajor;
    uint16_t vminor;
    uint16_t vpatch;
    uint16_t flags;
} kfft_info_t;

/* incs/kfft_types.h:82-94 */
enum kfft_ret_flags {
    KFFT_RET_SUCCESS = 0x0000,
    KFFT_RET_ALLOC_FAIL = 0x0001,
    KFFT_RET_BUFFER_FAIL = 0x0002,
    KFFT_RET_FREE_NULL = 0x0003,
    KFFT_RET_IMPROPER_PLAN = 0x0004,
    KFFT_RET_BADARGUMENTS = 0x0005,
    KFFT_RET_GPU_FAIL = 0x0006, /* GPU: a HIP call failed; kfft_gpu_last_error() has the text */
};

typedef unsigned kfft_return_t;

/* incs/kfft_types.h:100-103 */
typedef struct {
    uint8_t arch;
    uint32_t ext;
} kfft_simd_t;
