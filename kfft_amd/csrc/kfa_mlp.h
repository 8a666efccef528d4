// The planner's cost model (kfa_mlp.hip): nin -> H -> H -> 1 MLP, ReLU, FP32,
// trained with Adam on mean squared error, all on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include "kfa_engine.h"

namespace kfa {

class Mlp {
  public:
    Mlp(uint32_t nin, uint32_t hidden);
    bool init(uint32_t seed);
    uint32_t nin() const { return nin_; }
    uint32_t hidden() const { return h_; }
    uint32_t param_count() const { return np_; }
    bool set_params(const float* host); // also clears the Adam state
    bool get_params(float* host) const;
    bool get_grads(float* host) const;
    // device pointers: X [B x nin] row-major, y / t [B]
    hipError_t forward_dev(const float* X, float* y, uint32_t B, hipStream_t s);
    // gradients of the mean squared error into the gradient buffer; loss on the host if asked
    hipError_t gradient_dev(const float* X, const float* t, uint32_t B, float* loss, hipStream_t s);
    hipError_t train_step(const float* X, const float* t, uint32_t B, float lr, float* loss,
                          hipStream_t s);

  private:
    uint32_t nin_, h_, np_, cap_ = 0, step_ = 0;
    DevBuf params_, grads_, m_, v_;
    DevBuf z1_, a1_, z2_, a2_, d1_, d2_, y_, dy_, part_;
    bool ensure(uint32_t B);
    uint32_t off_b1() const { return h_ * nin_; }
    uint32_t off_w2() const { return h_ * nin_ + h_; }
    uint32_t off_b2() const { return off_w2() + h_ * h_; }
    uint32_t off_w3() const { return off_b2() + h_; }
    uint32_t off_b3() const { return off_w3() + h_; }
};

} // namespace kfa
