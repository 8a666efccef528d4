/* 1D complex transform API (incs/kfft_cpx.h:15-152). */
#pragma once

#define MAX_FACTORS 32
#define MAX_ROOTS 32

/* Every plan starts with this object (incs/kfft_cpx.h:21-23). */
typedef struct {
    kfft_pool_t* mmgr;
} kfft_object_t;

/* incs/kfft_cpx.h:29-41.  The GPU engine keeps its own Rader tables on the
 * device; here only `prime` is filled, for the primes the reference's
 * factoriser would register. */
typedef struct {
    uint32_t prime;
    uint32_t p, q;
    uint32_t* qidx;
    uint32_t* pidx;
    kfft_cpx* shuffle_twiddles;
    struct kfft_kstate* plan;
    struct kfft_kstate* plan_inv;
} kfft_plan_rader;

/* incs/kfft_cpx.h:46-65, plus the handle of the GPU plan at the end. */
typedef struct kfft_kstate {
    kfft_object_t object;
    uint32_t nfft;
    uint8_t level;
    uint32_t flags;
    uint8_t fac_count;
    uint32_t factors[2 * MAX_FACTORS];
    uint8_t prm_count;
    kfft_plan_rader primes[MAX_ROOTS];
    kfft_cpx* twiddles;
    void* gpu; /* GPU: device plan */
} kfft_plan_cpx;

KFFT_API kfft_plan_cpx* kfft_config_cpx(const uint32_t nfft, const uint32_t flags, kfft_pool_t* A,
                                        size_t* lenmem);
KFFT_API kfft_return_t kfft_eval_cpx(kfft_plan_cpx* plan, const kfft_cpx* fin, kfft_cpx* fout);

#define KFFT_PLAN_MMGR(X) (*((kfft_pool_t**)(X)))
#define KFFT_PLAN_MMGR_NULL(X) ((X != NULL) ? KFFT_PLAN_MMGR((X)) : NULL)
#define KFFT_PLAN_ALIGN(X) ((X != NULL) ? KFFT_PLAN_MMGR((X))->align : 0)
#define KFFT_PLAN_VEX(X) KFFT_PLAN_MMGR((X))->vex

#if defined(KFFT_DYNAPI_ENABLE)
typedef kfft_plan_cpx* (*kfft_callback_config_cpx)(const uint32_t nfft, const uint32_t flags,
                                                   kfft_pool_t* A, size_t* lenmem);
typedef kfft_return_t (*kfft_callback_eval_cpx)(kfft_plan_cpx* plan, const kfft_cpx* fin,
                                                kfft_cpx* fout);
#endif
