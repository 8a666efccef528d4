/*
 * Host SIMD information (reference simd/kfft_simd_check.h:1-66, same names,
 * values and signatures).  This library has no CPU SIMD backend: its
 * transforms run on the GPU, so kfft_simd_info (the build's CPU extensions)
 * reports the host architecture with no extensions, and kfft_simd_analize
 * reports what the host CPU supports, as the reference does.  Not included
 * by kfft.h (as in the reference).
 */
#pragma once
#include <stdbool.h>
#include <stdint.h>

#include "kfft_config.h"
#include "kfft_macro.h"
#include "kfft_types.h"

#define HW_ALIGN_X8664_SSE 16
#define HW_ALIGN_X8664_AVX 32
#define HW_ALIGN_ARM_NEON 16

enum {
    HW_ARCH_UKNW = 0,
    HW_ARCH_X86 = 1u << 0,
    HW_ARCH_X64 = 1u << 1,
    HW_ARCH_ARM = 1u << 2,
};

/* x86 extension bits of kfft_simd_t::ext */
#define HW_MMX (1u << 0)
#define HW_x64 (1u << 1)
#define HW_ABM (1u << 2)
#define HW_RDRAND (1u << 3)
#define HW_BMI1 (1u << 4)
#define HW_BMI2 (1u << 5)
#define HW_ADX (1u << 6)
#define HW_PREFETCHWT1 (1u << 7)
#define HW_MPX (1u << 8)
#define HW_SSE (1u << 9)
#define HW_SSE2 (1u << 10)
#define HW_SSE3 (1u << 11)
#define HW_SSSE3 (1u << 12)
#define HW_SSE41 (1u << 13)
#define HW_SSE42 (1u << 14)
#define HW_SSE4a (1u << 15)
#define HW_AES (1u << 16)
#define HW_SHA (1u << 17)
#define HW_AVX (1u << 18)
#define HW_XOP (1u << 19)
#define HW_FMA3 (1u << 20)
#define HW_FMA4 (1u << 21)
#define HW_AVX2 (1u << 22)
#define HW_AVX512_F (1u << 23)
#define HW_AVX512_PF (1u << 24)
#define HW_AVX512_ER (1u << 25)
#define HW_AVX512_CD (1u << 26)
#define HW_AVX512_VL (1u << 27)
#define HW_AVX512_BW (1u << 28)
#define HW_AVX512_DQ (1u << 29)
#define HW_AVX512_IFMA (1u << 30)
#define HW_AVX512_VBMI (1u << 31)

#ifdef __cplusplus
extern "C" {
#endif

/* replaces simd/kfft_simd_check.h:58-59: host CPU architecture and extensions */
KFFT_API kfft_simd_t kfft_simd_analize(void);
/* replaces simd/kfft_simd_check.h:65-66: CPU extensions the library was built
 * to use (none: the transforms run on the GPU) */
KFFT_API kfft_simd_t kfft_simd_info(void);

#ifdef __cplusplus
}
#endif

static inline bool kfft_simd_check(const kfft_simd_t S, uint32_t flag) { return (S.ext & flag) != 0; }

/* buffer alignment the extensions of S ask for (0 when none) */
static inline uint8_t kfft_simd_align(const kfft_simd_t S) {
    if (S.arch != HW_ARCH_X86 && S.arch != HW_ARCH_X64)
        return 0;
    if (kfft_simd_check(S, HW_AVX2 | HW_AVX))
        return HW_ALIGN_X8664_AVX;
    if (kfft_simd_check(S, HW_SSE | HW_SSE2))
        return HW_ALIGN_X8664_SSE;
    return 0;
}
