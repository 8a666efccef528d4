/*
 * Tiny API: one plan handle for every plan kind, over this library's plans.
 * Same interface as the reference's header-only kfft_tiny.h (kfft_tiny.h:22-275:
 * the ktiny_plan_type values, the kfft_plan_t layout {state, type}, the
 * argument structs and the kfft_configuration / kfft_evaluation /
 * kfft_convolution / kfft_finalization macros), written for this library.
 * The reference's own header does not compile in its tree without the sparse
 * extension (kfft_tiny_eval leaves its switch open when KFFT_SPARSE_ENABLE is
 * undefined, and that extension's headers are absent); this one has no sparse
 * plan kinds.
 *
 * Buffers may be host or device memory, as for the underlying entry points.
 */
#pragma once

#include "kfft.h"

#ifdef __cplusplus
extern "C" {
#endif

enum ktiny_plan_type {
    KFFT_PLAN_COMPLEX = 0x0000,        /* 1D complex */
    KFFT_PLAN_SCALAR = 0x0001,         /* 1D real (forward or KFFT_FLAG_INVERSE) */
    KFFT_PLAN_COMPLEX_2D = 0x0002,     /* 2D complex */
    KFFT_PLAN_SCALAR_2D = 0x0003,      /* 2D real */
    KFFT_PLAN_COMPLEX_CONV = 0x0006,   /* 1D complex circular convolution */
    KFFT_PLAN_SCALAR_CONV = 0x0007,    /* 1D real circular convolution */
    KFFT_PLAN_COMPLEX_CONV2D = 0x0008, /* 2D complex circular convolution */
    KFFT_PLAN_SCALAR_CONV2D = 0x0009,  /* 2D real circular convolution */
};

/* a plan of any kind: `state` is the kfft_plan_* pointer, `type` its kind */
typedef struct {
    uintptr_t state;
    int type;
} kfft_plan_t;
typedef kfft_plan_t* kfft_plan;

#define KTINY_CAST(type, X) (kfft_plan_##type*)((X)->state)
#define KTINY_FLAGS(type, X) (KTINY_CAST(type, (X)))->flags
#define KTINY_CPX(X) (kfft_cpx*)(X)
#define KTINY_SCLR(X) (kfft_scalar*)(X)

/* configuration arguments: 1D length, or 2D sizes */
typedef struct {
    uint32_t nfft;
} kfft_args_norm;
typedef struct {
    uint32_t x;
    uint32_t y;
} kfft_args_2d;

/* T: ktiny_plan_type, F: flags (KFFT_FLAG_RENEW is ignored), A: kfft_args_* pointer */
#define kfft_configuration(T, F, A) kfft_tiny_config((T), (F), (uintptr_t)(A))
#define kfft_evaluation(P, Fi, Fo) kfft_tiny_eval(P, (kfft_scalar*)(Fi), (kfft_scalar*)(Fo))
#define kfft_convolution(P, Fi1, Fi2, Fo)                                                          \
    kfft_tiny_conv(P, (kfft_scalar*)(Fi1), (kfft_scalar*)(Fi2), (kfft_scalar*)(Fo))
/* releases the plan and sets P to NULL */
#define kfft_finalization(P) kfft_tiny_release(&(P))

static kfft_plan KFFT_UNUSED_FUNC kfft_tiny_config(unsigned type, uint32_t flags, uintptr_t args) {
    const kfft_args_norm* a1 = (const kfft_args_norm*)args;
    const kfft_args_2d* a2 = (const kfft_args_2d*)args;
    void* state = NULL;
    kfft_plan plan;
    if (!args)
        return NULL;
    flags &= ~(uint32_t)KFFT_FLAG_RENEW;
    switch (type) {
    case KFFT_PLAN_COMPLEX:
        state = kfft_config_cpx(a1->nfft, flags, NULL, NULL);
        break;
    case KFFT_PLAN_SCALAR:
        state = kfft_config_scalar(a1->nfft, flags, NULL, NULL);
        break;
    case KFFT_PLAN_COMPLEX_2D:
        state = kfft_config2_cpx(a2->x, a2->y, flags, NULL, NULL);
        break;
    case KFFT_PLAN_SCALAR_2D:
        state = kfft_config2_scalar(a2->x, a2->y, flags, NULL, NULL);
        break;
    case KFFT_PLAN_COMPLEX_CONV:
        state = kfft_config_conv_cpx(a1->nfft, flags, NULL, NULL);
        break;
    case KFFT_PLAN_SCALAR_CONV:
        state = kfft_config_conv_scalar(a1->nfft, flags, NULL, NULL);
        break;
    case KFFT_PLAN_COMPLEX_CONV2D:
        state = kfft_config2_conv_cpx(a2->x, a2->y, flags, NULL, NULL);
        break;
    case KFFT_PLAN_SCALAR_CONV2D:
        state = kfft_config2_conv_scalar(a2->x, a2->y, flags, NULL, NULL);
        break;
    default:
        return NULL;
    }
    if (!state)
        return NULL;
    plan = (kfft_plan)kfft_malloc(sizeof(kfft_plan_t));
    if (!plan) {
        kfft_cleanup(state);
        return NULL;
    }
    plan->state = (uintptr_t)state;
    plan->type = (int)type;
    return plan;
}

static kfft_return_t KFFT_UNUSED_FUNC kfft_tiny_eval(const kfft_plan plan, const kfft_scalar* fin,
                                                     kfft_scalar* fout) {
    if (!plan)
        return KFFT_RET_IMPROPER_PLAN;
    switch (plan->type) {
    case KFFT_PLAN_COMPLEX:
        return kfft_eval_cpx(KTINY_CAST(cpx, plan), KTINY_CPX(fin), KTINY_CPX(fout));
    case KFFT_PLAN_SCALAR:
        if (KTINY_FLAGS(sclr, plan) & KFFT_FLAG_INVERSE)
            return kfft_evali_scalar(KTINY_CAST(sclr, plan), KTINY_CPX(fin), fout);
        return kfft_eval_scalar(KTINY_CAST(sclr, plan), fin, KTINY_CPX(fout));
    case KFFT_PLAN_COMPLEX_2D:
        return kfft_eval2_cpx(KTINY_CAST(c2d, plan), KTINY_CPX(fin), KTINY_CPX(fout));
    case KFFT_PLAN_SCALAR_2D:
        if (KTINY_FLAGS(s2d, plan) & KFFT_FLAG_INVERSE)
            return kfft_evali2_scalar(KTINY_CAST(s2d, plan), KTINY_CPX(fin), fout);
        return kfft_eval2_scalar(KTINY_CAST(s2d, plan), fin, KTINY_CPX(fout));
    default:
        return KFFT_RET_IMPROPER_PLAN;
    }
}

static kfft_return_t KFFT_UNUSED_FUNC kfft_tiny_conv(const kfft_plan plan, const kfft_scalar* fin1,
                                                     const kfft_scalar* fin2, kfft_scalar* fout) {
    if (!plan)
        return KFFT_RET_IMPROPER_PLAN;
    switch (plan->type) {
    case KFFT_PLAN_COMPLEX_CONV:
        return kfft_eval_conv_cpx(KTINY_CAST(ccnv, plan), KTINY_CPX(fin1), KTINY_CPX(fin2),
                                  KTINY_CPX(fout));
    case KFFT_PLAN_SCALAR_CONV:
        return kfft_eval_conv_scalar(KTINY_CAST(scnv, plan), fin1, fin2, fout);
    case KFFT_PLAN_COMPLEX_CONV2D:
        return kfft_eval2_conv_cpx(KTINY_CAST(c2cnv, plan), KTINY_CPX(fin1), KTINY_CPX(fin2),
                                   KTINY_CPX(fout));
    case KFFT_PLAN_SCALAR_CONV2D:
        return kfft_eval2_conv_scalar(KTINY_CAST(s2cnv, plan), fin1, fin2, fout);
    default:
        return KFFT_RET_IMPROPER_PLAN;
    }
}

static void KFFT_UNUSED_FUNC kfft_tiny_release(kfft_plan* plan) {
    if (!plan || !*plan)
        return;
    kfft_cleanup((void*)((*plan)->state));
    kfft_free(*plan);
}

#ifdef __cplusplus
}
#endif
