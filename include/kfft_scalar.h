/* 1D real-input ("scalar") transform API (incs/kfft_scalar.h:15-103). */
#pragma once

typedef struct kfft_state {
    kfft_object_t object;
    uint32_t nfft;
    uint32_t flags;
    kfft_plan_cpx* basis;
    kfft_cpx* super_twiddles;
    void* gpu; /* GPU: device plan */
} kfft_plan_sclr;

KFFT_API kfft_plan_sclr* kfft_config_scalar(const uint32_t nfft, const uint32_t flags,
                                            kfft_pool_t* A, size_t* lenmem);
KFFT_API kfft_return_t kfft_eval_scalar(kfft_plan_sclr* plan, const kfft_scalar* fin,
                                        kfft_cpx* fout);
KFFT_API kfft_return_t kfft_evali_scalar(kfft_plan_sclr* plan, const kfft_cpx* fin,
                                         kfft_scalar* fout);
/* incs/kfft_scalar.h:74-103: variants taking a caller temporary (ignored here) */
KFFT_API kfft_return_t kfft_eval_scalar_internal(kfft_plan_sclr* plan, const kfft_scalar* fin,
                                                 kfft_cpx* fout, kfft_cpx* ftmp);
KFFT_API kfft_return_t kfft_evali_scalar_internal(kfft_plan_sclr* plan, const kfft_cpx* fin,
                                                  kfft_scalar* fout, kfft_cpx* ftmp);

#if defined(KFFT_DYNAPI_ENABLE)
typedef kfft_plan_sclr* (*kfft_callback_config_scalar)(const uint32_t nfft, const uint32_t flags,
                                                       kfft_pool_t* A, size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_scalar)(kfft_plan_sclr* plan, const kfft_scalar* fin,
                                                   kfft_cpx* fout);
typedef kfft_return_t (*kfft_callback_evali_scalar)(kfft_plan_sclr* plan, const kfft_cpx* fin,
                                                    kfft_scalar* fout);
#endif
