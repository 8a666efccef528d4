/* Public math helpers used by the extensions (incs/kfft_math.h:251-350). */
#pragma once

KFFT_API void kfft_math_hadamard_cpx(kfft_cpx* Fout, const kfft_cpx* Fin, uint32_t size);
KFFT_API void kfft_math_hadamard_scalar(kfft_scalar* Fout, const kfft_scalar* Fin, uint32_t size);
KFFT_API void kfft_math_transpose_cpx(const kfft_cpx* Fin, kfft_cpx* Fout, const uint32_t x,
                                      const uint32_t y);
KFFT_API void kfft_math_transpose_scalar(const kfft_scalar* Fin, kfft_scalar* Fout,
                                         const uint32_t x, const uint32_t y);
/* replaces incs/kfft_math.h:303-304: in-place transpose, the result of
 * kfft_math_transpose_cpx written over the input */
KFFT_API void kfft_math_transpose_ip_cpx(kfft_cpx* Fin, const uint32_t x, const uint32_t y);
/* replaces incs/kfft_math.h:317-318 */
KFFT_API void kfft_math_transpose_ip_scalar(kfft_scalar* Fin, const uint32_t x, const uint32_t y);
/* replaces incs/kfft_math.h:328-329: Fout[i] = |Fin[i]| */
KFFT_API void kfft_math_magnitude(const kfft_cpx* Fin, kfft_scalar* Fout, uint32_t size);
/* replaces incs/kfft_math.h:339-340: |Fin[i]| into the first `size` reals of Fin */
KFFT_API void kfft_math_magnitude_ip(kfft_cpx* Fin, uint32_t size);
/* replaces incs/kfft_math.h:349-350: the mean of `size` reals */
KFFT_API kfft_scalar kfft_math_average(const kfft_scalar* Fin, uint32_t size);
