/* fftshift of 1D buffers (incs/kfft_shift.h:16-27). */
#pragma once

KFFT_API void kfft_shift_cpx(kfft_cpx* buf, const uint32_t size, const bool is_inverse,
                             kfft_pool_t* mmgr);
KFFT_API void kfft_shift_scalar(kfft_scalar* buf, const uint32_t size, const bool is_inverse,
                                kfft_pool_t* mmgr);
