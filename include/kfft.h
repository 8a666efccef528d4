/*
 * libkfft-compatible API of the MI355X (gfx950) transform engine.
 *
 * Restates the public declarations of the reference's kfft.h (kfft.h:63-100)
 * with the same names, types, flag values, return codes and plan-memory
 * (pool / lenmem) semantics, so that code written against libkfft 0.8.1
 * compiles and links unchanged.  Transforms execute on the GPU through
 * hand-written HIP kernels; kfft_gpu.h adds batched, asynchronous and
 * device-memory entry points.
 */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "kfft_config.h"

#ifdef __cplusplus
extern "C" {
#endif

#include "kfft_macro.h"
#include "kfft_types.h"
#include "kfft_pool.h"
#include "kfft_ext.h"
#include "kfft_cpx.h"
#include "kfft_scalar.h"
#include "kfft_math.h"
#include "kfft_shift.h"
#include "kfft_cpx2.h"
#include "kfft_scalar2.h"
#include "kfft_conv.h"
#include "kfft_gpu.h"

#ifdef __cplusplus
}
#endif
