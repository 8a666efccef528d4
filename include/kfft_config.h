/*
 * Build configuration of the MI355X (gfx950) kfft drop-in.
 *
 * Macro names follow the reference's generated configuration header
 * (config.h.in:1-65) so that code testing them keeps compiling.  Values
 * describe this library: API version 0.8.1, Rader on with the reference's
 * limits, all three extensions, no SIMD/OpenMP/trace host paths (the
 * transforms run on the GPU).
 */
#pragma once

#define KFFT_MAJOR_VERSION 0
#define KFFT_MINOR_VERSION 8
#define KFFT_PATCH_VERSION 1

#define KFFT_DYNAPI_ENABLE
#define KFFT_RADER_ALGO
#define KFFT_RADER_LIMIT 11
#define KFFT_PLAN_LEVEL 7
#define KFFT_BFLY_LEVEL 5

#define KFFT_2D_ENABLE
#define KFFT_CONV_ENABLE
#define KFFT_CONV2_ENABLE
#define KFFT_CONV2D_ENABLE

/* marks this implementation: transforms execute on an AMD GPU through HIP */
#define KFFT_GPU_HIP 1
