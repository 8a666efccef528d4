/* Export and utility macros (names as in the reference, incs/kfft_macro.h:101-137). */
#pragma once

#if defined(KFFT_EXPORTS)
#define KFFT_API __attribute__((visibility("default")))
#else
#define KFFT_API
#endif

#define KFFT_UNUSED_VAR(X) (void)(X)

#ifndef __likely__
#define __likely__(x) __builtin_expect(!!(x), 1)
#endif
#ifndef __unlikely__
#define __unlikely__(x) __builtin_expect(!!(x), 0)
#endif
