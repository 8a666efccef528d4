/* 1D and 2D circular convolution API (ext/conv/kfft_cpx_conv.h:5-21,
 * ext/conv/kfft_scalar_conv.h:5-19, ext/conv2d/kfft_cpx_conv.h:5-23,
 * ext/conv2d/kfft_scalar_conv.h:5-23). */
#pragma once

typedef struct {
    kfft_object_t object;
    uint32_t nfft;
    uint32_t flags;
    kfft_plan_cpx* plan_fwd;
    kfft_plan_cpx* plan_inv;
    void* gpu;
} kfft_plan_ccnv;

#define KFFT_CHECK_FLAGS_CCNV(X) (KFFT_CHECK_FLAGS(X) & (~KFFT_FLAG_INVERSE))

typedef struct {
    kfft_object_t object;
    uint32_t nfft;
    uint32_t flags;
    kfft_plan_sclr* plan_fwd;
    kfft_plan_sclr* plan_inv;
    void* gpu;
} kfft_plan_scnv;

typedef struct {
    kfft_object_t object;
    uint32_t nfft;
    uint32_t x;
    uint32_t y;
    uint32_t flags;
    kfft_plan_c2d* plan_fwd;
    kfft_plan_c2d* plan_inv;
    void* gpu;
} kfft_plan_c2cnv;

typedef struct {
    kfft_object_t object;
    uint32_t nfft;
    uint32_t x;
    uint32_t y;
    uint32_t flags;
    kfft_plan_s2d* plan_fwd;
    kfft_plan_s2d* plan_inv;
    void* gpu;
} kfft_plan_s2cnv;

KFFT_API kfft_plan_ccnv* kfft_config_conv_cpx(const uint32_t nfft, const uint32_t flags,
                                              kfft_pool_t* A, size_t* lenmem);
KFFT_API kfft_return_t kfft_eval_conv_cpx(kfft_plan_ccnv* plan, const kfft_cpx* fin_A,
                                          const kfft_cpx* fin_B, kfft_cpx* fout);
KFFT_API kfft_plan_scnv* kfft_config_conv_scalar(const uint32_t nfft, const uint32_t flags,
                                                 kfft_pool_t* A, size_t* lenmem);
KFFT_API kfft_return_t kfft_eval_conv_scalar(kfft_plan_scnv* plan, const kfft_scalar* fin_A,
                                             const kfft_scalar* fin_B, kfft_scalar* fout);
KFFT_API kfft_plan_c2cnv* kfft_config2_conv_cpx(const uint32_t x, const uint32_t y,
                                                const uint32_t flags, kfft_pool_t* A,
                                                size_t* lenmem);
KFFT_API kfft_return_t kfft_eval2_conv_cpx(kfft_plan_c2cnv* plan, const kfft_cpx* fin_A,
                                           const kfft_cpx* fin_B, kfft_cpx* fout);
KFFT_API kfft_plan_s2cnv* kfft_config2_conv_scalar(const uint32_t x, const uint32_t y,
                                                   const uint32_t flags, kfft_pool_t* A,
                                                   size_t* lenmem);
KFFT_API kfft_return_t kfft_eval2_conv_scalar(kfft_plan_s2cnv* plan, const kfft_scalar* fin_A,
                                              const kfft_scalar* fin_B, kfft_scalar* fout);

#if defined(KFFT_DYNAPI_ENABLE)
typedef kfft_plan_ccnv* (*kfft_callback_config_conv_cpx)(const uint32_t nfft, const uint32_t flags,
                                                         kfft_pool_t* A, size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_conv_cpx)(kfft_plan_ccnv* plan, const kfft_cpx* fin_A,
                                                     const kfft_cpx* fin_B, kfft_cpx* fout);
typedef kfft_plan_scnv* (*kfft_callback_config_conv_scalar)(const uint32_t nfft,
                                                            const uint32_t flags, kfft_pool_t* A,
                                                            size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_conv_scalar)(kfft_plan_scnv* plan,
                                                        const kfft_scalar* fin_A,
                                                        const kfft_scalar* fin_B,
                                                        kfft_scalar* fout);
typedef kfft_plan_c2cnv* (*kfft_callback_config_conv2_cpx)(const uint32_t x, const uint32_t y,
                                                           const uint32_t flags, kfft_pool_t* A,
                                                           size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_conv2_cpx)(kfft_plan_c2cnv* plan,
                                                      const kfft_cpx* fin_A,
                                                      const kfft_cpx* fin_B, kfft_cpx* fout);
typedef kfft_plan_s2cnv* (*kfft_callback_config_conv2_scalar)(const uint32_t x, const uint32_t y,
                                                              const uint32_t flags,
                                                              kfft_pool_t* A, size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_conv2_scalar)(kfft_plan_s2cnv* plan,
                                                         const kfft_scalar* fin_A,
                                                         const kfft_scalar* fin_B,
                                                         kfft_scalar* fout);
#endif
