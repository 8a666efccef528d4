/*
 * Plan arena (the reference's kfft_pool_t, incs/kfft_pool.h:14-75).  One
 * bump allocator holds a plan tree; device memory that plans own is
 * registered with the pool and released when the pool is freed or cleared.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

typedef struct {
    size_t allocated;
    uint8_t align;
    kfft_simd_t vex;

    uint8_t* head;
    uint8_t* tail;
    uint8_t* cur;

    void* gpu; /* GPU: registry of device objects owned by plans in this pool */

    uint8_t area[1];
} kfft_pool_t;

KFFT_API kfft_pool_t* kfft_pool_create(const size_t size);
KFFT_API void* kfft_pool_alloc(kfft_pool_t* A, const size_t nmem);
KFFT_API size_t kfft_pool_empty(const kfft_pool_t* A);
KFFT_API void kfft_pool_zmem(const kfft_pool_t* A, void* ptr, const size_t size);
KFFT_API void kfft_pool_clear(kfft_pool_t* A);
KFFT_API void kfft_pool_free(kfft_pool_t* A);

#define kfft_pool_free_and_null(A)                                                                 \
    do {                                                                                           \
        kfft_pool_free(A);                                                                         \
        A = NULL;                                                                                  \
    } while (0)
