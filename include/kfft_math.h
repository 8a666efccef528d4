/* Public math helpers used by the extensions (incs/kfft_math.h:251-350). */
#pragma once

KFFT_API void kfft_math_hadamard_cpx(kfft_cpx* Fout, const kfft_cpx* Fin, uint32_t size);
KFFT_API void kfft_math_hadamard_scalar(kfft_scalar* Fout, const kfft_scalar* Fin, uint32_t size);
KFFT_API void kfft_math_transpose_cpx(const kfft_cpx* Fin, kfft_cpx* Fout, const uint32_t x,
                                      const uint32_t y);
KFFT_API void kfft_math_transpose_scalar(const kfft_scalar* Fin, kfft_scalar* Fout,
                                         const uint32_t x, const uint32_t y);
