/*
 * GPU additions of the MI355X drop-in.  The reference has no counterpart: it
 * has neither a batch API nor device memory.  Everything here is additive;
 * the reference's entry points keep their signatures and semantics.
 *
 * Pointers: every eval entry point of kfft.h accepts host or device memory.
 * Device (or managed) buffers are transformed in place on the GPU; host
 * buffers are staged through device memory (correct but PCIe-bound).
 * The *_async variants take device pointers only and return without
 * synchronising; `stream` is a hipStream_t (NULL: the library's stream).
 */
#pragma once

/* device memory and synchronisation */
KFFT_API void* kfft_gpu_malloc(size_t bytes);
KFFT_API void kfft_gpu_free(void* ptr);
KFFT_API kfft_return_t kfft_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes);
KFFT_API kfft_return_t kfft_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes);
KFFT_API kfft_return_t kfft_gpu_synchronize(void);
KFFT_API void* kfft_gpu_stream(void);
KFFT_API const char* kfft_gpu_last_error(void);
/* non-zero when an asynchronous launch of this complex plan reported a failed
 * dependency wait since the last call (synchronises the library stream) */
KFFT_API int kfft_gpu_plan_error(const kfft_plan_cpx* plan);
/* one line describing how a complex, scalar or 2D plan is executed */
KFFT_API const char* kfft_gpu_plan_describe(const void* plan);

/* `batch` contiguous transforms: complex in/out distance nfft */
KFFT_API kfft_return_t kfft_eval_cpx_batch(kfft_plan_cpx* plan, const kfft_cpx* fin,
                                           kfft_cpx* fout, uint32_t batch);
KFFT_API kfft_return_t kfft_eval_cpx_batch_async(kfft_plan_cpx* plan, const kfft_cpx* fin,
                                                 kfft_cpx* fout, uint32_t batch, void* stream);

/* real input: input distance nfft reals; output distance kfft_scalar_out_len(plan) */
KFFT_API uint32_t kfft_scalar_out_len(const kfft_plan_sclr* plan);
KFFT_API kfft_return_t kfft_eval_scalar_batch(kfft_plan_sclr* plan, const kfft_scalar* fin,
                                              kfft_cpx* fout, uint32_t batch);
KFFT_API kfft_return_t kfft_evali_scalar_batch(kfft_plan_sclr* plan, const kfft_cpx* fin,
                                               kfft_scalar* fout, uint32_t batch);
KFFT_API kfft_return_t kfft_eval_scalar_batch_async(kfft_plan_sclr* plan, const kfft_scalar* fin,
                                                    kfft_cpx* fout, uint32_t batch, void* stream);
KFFT_API kfft_return_t kfft_evali_scalar_batch_async(kfft_plan_sclr* plan, const kfft_cpx* fin,
                                                     kfft_scalar* fout, uint32_t batch,
                                                     void* stream);

/* 2D, asynchronous on device pointers */
KFFT_API kfft_return_t kfft_eval2_cpx_async(kfft_plan_c2d* plan, const kfft_cpx* fin,
                                            kfft_cpx* fout, void* stream);

/*
 * Ragged batch (the batching form): `count` transforms of uneven lengths run
 * as one persistent launch fed by device work queues, replayed as a captured
 * hipGraph.  kind[i]: KFFT_RAGGED_COMPLEX (complex, nfft[i] outputs, forward or
 * KFFT_FLAG_INVERSE as for kfft_eval_cpx) or KFFT_RAGGED_REAL (real-input
 * forward: nfft/2+1 outputs for even nfft, nfft for odd, as kfft_eval_scalar).
 * Offsets are in bytes from the base pointers given at evaluation; NULL
 * offsets select the packed layout (transforms in order, each starting on a
 * 256-byte boundary), which kfft_ragged_layout reports.  Data is kfft_cpx /
 * kfft_scalar of the build.  Device pointers only.
 */
#define KFFT_RAGGED_COMPLEX 0
#define KFFT_RAGGED_REAL 1
typedef struct kfft_plan_ragged kfft_plan_ragged;
KFFT_API kfft_plan_ragged* kfft_config_ragged(uint32_t count, const uint32_t* nfft,
                                              const uint8_t* kind, const uint64_t* in_off,
                                              const uint64_t* out_off, uint32_t flags);
/* offsets (count entries each, may be NULL) and the bytes the buffers must span */
KFFT_API kfft_return_t kfft_ragged_layout(const kfft_plan_ragged* plan, uint64_t* in_off,
                                          uint64_t* out_off, uint64_t* in_bytes,
                                          uint64_t* out_bytes);
KFFT_API kfft_return_t kfft_eval_ragged(kfft_plan_ragged* plan, const void* in, void* out);
KFFT_API kfft_return_t kfft_eval_ragged_async(kfft_plan_ragged* plan, const void* in, void* out,
                                              void* stream);
/* non-zero when the last evaluation's dependency waits timed out (synchronises) */
KFFT_API int kfft_ragged_error(kfft_plan_ragged* plan);
KFFT_API const char* kfft_ragged_describe(const kfft_plan_ragged* plan);
KFFT_API void kfft_release_ragged(kfft_plan_ragged* plan);
