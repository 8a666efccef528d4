/* Export and utility macros (names as in the reference, incs/kfft_macro.h:1-165).
 * The library is built for Linux with GCC-compatible compilers only (ROCm), so
 * only the platform macros of those targets are defined. */
#pragma once

#if defined(KFFT_EXPORTS)
#define KFFT_API __attribute__((visibility("default")))
#else
#define KFFT_API
#endif

/* platform and compiler (incs/kfft_macro.h:5-80) */
#if defined(__linux) || defined(__linux__)
#define KFFT_OS_LINUX
#define KFFT_OS_POSIX
#endif
#if defined(__GNUC__)
#define KFFT_CC_GNU
#if defined(__clang__)
#define KFFT_CC_CLANG
#else
#define KFFT_CC_GCC
#endif
#endif

#define KFFT_STRINGIZE_DETAIL(X) #X
#define KFFT_STRINGIZE(X) KFFT_STRINGIZE_DETAIL(X)
#define KFFT_PRAGMA(X) _Pragma(#X)

#define KFFT_UNUSED_VAR(X) (void)(X)
/* marks the header-only helpers (kfft_tiny.h) that a unit may leave unused */
#define KFFT_UNUSED_FUNC __attribute__((__unused__))

#ifndef __likely__
#define __likely__(x) __builtin_expect(!!(x), 1)
#endif
#ifndef __unlikely__
#define __unlikely__(x) __builtin_expect(!!(x), 0)
#endif
