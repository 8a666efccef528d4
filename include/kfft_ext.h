/* Service API (incs/kfft_ext.h:16-66). */
#pragma once

KFFT_API void kfft_info(kfft_info_t* info);
KFFT_API uint32_t kfft_next_fast_size(uint32_t n);
KFFT_API void* kfft_malloc(uint32_t sz);
KFFT_API void kfft_free_null(void** mem);
#define kfft_free(X) (kfft_free_null((void**)(&(X))))
KFFT_API kfft_return_t kfft_cleanup(void* mem);
#define kfft_release(X)                                                                            \
    do {                                                                                           \
        if (kfft_cleanup((void*)(X)) == KFFT_RET_SUCCESS)                                          \
            (X) = NULL;                                                                            \
    } while (0)
KFFT_API const char* kfft_strerr(const kfft_return_t code);

#if defined(KFFT_DYNAPI_ENABLE)
typedef void (*kfft_callback_info)(kfft_info_t* info);
typedef uint32_t (*kfft_callback_next_fast_size)(uint32_t n);
typedef void* (*kfft_callback_malloc)(uint32_t sz);
typedef void (*kfft_callback_free_null)(void** mem);
typedef kfft_return_t (*kfft_callback_cleanup)(void* mem);
typedef const char* (*kfft_callback_strerr)(const kfft_return_t code);
#endif
