// The planner's cost model: a small MLP (nin -> H -> H -> 1, ReLU, FP32)
// trained and evaluated by these kernels.  It predicts ln(microseconds) of
// one length class under one plan choice; it never touches transform data.
//
// Parameters live in one device buffer, in the order W1 [H x nin], b1 [H],
// W2 [H x H], b2 [H], W3 [1 x H], b3 [1]; gradients and the two Adam moments
// use the same layout.  Loss: mean squared error over the batch.
//   forward   Z1 = X W1^T + b1, A1 = relu(Z1), Z2 = A1 W2^T + b2,
//             A2 = relu(Z2), y = A2 W3^T + b3
//   backward  dy = 2 (y - t) / B; dW3 = dy^T A2, db3 = sum dy;
//             dZ2 = (dy W3) * [Z2 > 0]; dW2 = dZ2^T A1; db2 = sum dZ2;
//             dZ1 = (dZ2 W2) * [Z1 > 0]; dW1 = dZ1^T X; db1 = sum dZ1
// Every reduction runs in a fixed order (one thread per output, sequential
// over the batch), so results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "kfa_engine.h"
#include "kfa_mlp.h"

namespace kfa {

// Z[b][o] = bias[o] + sum_i W[o][i] X[b][i];  A = relu(Z) (or Z on the last layer)
__global__ void k_dense_fwd(const float* X, const float* W, const float* bias, float* Z, float* A,
                            uint32_t B, uint32_t nin, uint32_t nout, int relu) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nout)
        return;
    const uint32_t b = idx / nout, o = idx - b * nout;
    const float* x = X + (size_t)b * nin;
    const float* w = W + (size_t)o * nin;
    float z = bias[o];
    for (uint32_t i = 0; i < nin; ++i)
        z = fmaf(w[i], x[i], z);
    if (Z)
        Z[idx] = z;
    A[idx] = relu ? fmaxf(z, 0.0f) : z;
}

// dy[b] = 2 (y[b] - t[b]) / B;  per-block partial sums of (y - t)^2
__global__ void k_mse_grad(const float* y, const float* t, float* dy, float* partial, uint32_t B) {
    __shared__ float red[256];
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (b < B) {
        const float d = y[b] - t[b];
        dy[b] = 2.0f * d / (float)B;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = red[0];
}

// dW[o][i] = sum_b dZ[b][o] X[b][i] (i < nin); db[o] = sum_b dZ[b][o] (i == nin)
__global__ void k_dense_bwd_w(const float* dZ, const float* X, float* dW, float* db, uint32_t B,
                              uint32_t nin, uint32_t nout) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nout * (nin + 1))
        return;
    const uint32_t o = idx / (nin + 1), i = idx - o * (nin + 1);
    float s = 0.0f;
    if (i < nin) {
        for (uint32_t b = 0; b < B; ++b)
            s = fmaf(dZ[(size_t)b * nout + o], X[(size_t)b * nin + i], s);
        dW[(size_t)o * nin + i] = s;
    } else {
        for (uint32_t b = 0; b < B; ++b)
            s += dZ[(size_t)b * nout + o];
        db[o] = s;
    }
}

// dZprev[b][i] = (sum_o dZ[b][o] W[o][i]) * [Zprev[b][i] > 0]
__global__ void k_dense_bwd_x(const float* dZ, const float* W, const float* Zprev, float* dZprev,
                              uint32_t B, uint32_t nin, uint32_t nout) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * nin)
        return;
    const uint32_t b = idx / nin, i = idx - b * nin;
    float s = 0.0f;
    for (uint32_t o = 0; o < nout; ++o)
        s = fmaf(dZ[(size_t)b * nout + o], W[(size_t)o * nin + i], s);
    dZprev[idx] = Zprev[idx] > 0.0f ? s : 0.0f;
}

// Adam: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr m^ / (sqrt(v^) + eps)
__global__ void k_adam(float* p, const float* g, float* m, float* v, uint32_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static unsigned blocks(uint64_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ----------------------------------------------------------------- host side
Mlp::Mlp(uint32_t nin, uint32_t hidden) : nin_(nin), h_(hidden) {
    np_ = h_ * nin_ + h_ + h_ * h_ + h_ + h_ + 1;
}

bool Mlp::init(uint32_t seed) {
    if (!params_.alloc((size_t)np_ * 4) || !grads_.alloc((size_t)np_ * 4) ||
        !m_.alloc((size_t)np_ * 4) || !v_.alloc((size_t)np_ * 4))
        return false;
    // He-uniform weights from a fixed 64-bit LCG, zero biases
    std::vector<float> p(np_, 0.0f);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((s >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    const float a1 = std::sqrt(6.0f / nin_), a2 = std::sqrt(6.0f / h_);
    for (uint32_t i = 0; i < h_ * nin_; ++i)
        p[i] = a1 * uni();
    for (uint32_t i = 0; i < h_ * h_; ++i)
        p[off_w2() + i] = a2 * uni();
    for (uint32_t i = 0; i < h_; ++i)
        p[off_w3() + i] = a2 * uni();
    return set_params(p.data());
}

bool Mlp::set_params(const float* host) {
    step_ = 0;
    return hip_ok(hipMemcpy(params_.p, host, (size_t)np_ * 4, hipMemcpyHostToDevice), "mlp set") &&
           hip_ok(hipMemset(m_.p, 0, (size_t)np_ * 4), "mlp adam m") &&
           hip_ok(hipMemset(v_.p, 0, (size_t)np_ * 4), "mlp adam v");
}

bool Mlp::get_params(float* host) const {
    return hip_ok(hipMemcpy(host, params_.p, (size_t)np_ * 4, hipMemcpyDeviceToHost), "mlp get");
}

bool Mlp::ensure(uint32_t B) {
    if (B <= cap_)
        return true;
    const size_t hb = (size_t)B * h_ * 4;
    if (!z1_.alloc(hb) || !a1_.alloc(hb) || !z2_.alloc(hb) || !a2_.alloc(hb) || !d2_.alloc(hb) ||
        !d1_.alloc(hb) || !y_.alloc((size_t)B * 4) || !dy_.alloc((size_t)B * 4) ||
        !part_.alloc((size_t)blocks(B) * 4))
        return false;
    cap_ = B;
    return true;
}

hipError_t Mlp::forward_dev(const float* X, float* y, uint32_t B, hipStream_t s) {
    if (!ensure(B))
        return hipErrorOutOfMemory;
    const float* P = (const float*)params_.p;
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)B * h_)), dim3(256), 0, s, X, P,
                       P + off_b1(), (float*)z1_.p, (float*)a1_.p, B, nin_, h_, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)B * h_)), dim3(256), 0, s,
                       (const float*)a1_.p, P + off_w2(), P + off_b2(), (float*)z2_.p,
                       (float*)a2_.p, B, h_, h_, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(B)), dim3(256), 0, s, (const float*)a2_.p,
                       P + off_w3(), P + off_b3(), (float*)nullptr, y, B, h_, 1u, 0);
    return hipGetLastError();
}

hipError_t Mlp::gradient_dev(const float* X, const float* t, uint32_t B, float* loss,
                             hipStream_t s) {
    if (!ensure(B))
        return hipErrorOutOfMemory;
    hipError_t e = forward_dev(X, (float*)y_.p, B, s);
    if (e != hipSuccess)
        return e;
    const float* P = (const float*)params_.p;
    float* G = (float*)grads_.p;
    const unsigned nb = blocks(B);
    hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, s, (const float*)y_.p, t,
                       (float*)dy_.p, (float*)part_.p, B);
    // layer 3: dZ3 = dy [B x 1]
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks(h_ + 1)), dim3(256), 0, s, (const float*)dy_.p,
                       (const float*)a2_.p, G + off_w3(), G + off_b3(), B, h_, 1u);
    hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks((uint64_t)B * h_)), dim3(256), 0, s,
                       (const float*)dy_.p, P + off_w3(), (const float*)z2_.p, (float*)d2_.p, B,
                       h_, 1u);
    // layer 2
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)h_ * (h_ + 1))), dim3(256), 0, s,
                       (const float*)d2_.p, (const float*)a1_.p, G + off_w2(), G + off_b2(), B, h_,
                       h_);
    hipLaunchKernelGGL(k_dense_bwd_x, dim3(blocks((uint64_t)B * h_)), dim3(256), 0, s,
                       (const float*)d2_.p, P + off_w2(), (const float*)z1_.p, (float*)d1_.p, B,
                       h_, h_);
    // layer 1
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(blocks((uint64_t)h_ * (nin_ + 1))), dim3(256), 0, s,
                       (const float*)d1_.p, X, G, G + off_b1(), B, nin_, h_);
    e = hipGetLastError();
    if (e != hipSuccess || !loss)
        return e;
    std::vector<float> part(nb);
    e = hipMemcpyAsync(part.data(), part_.p, nb * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    double sum = 0;
    for (float v : part)
        sum += v;
    *loss = (float)(sum / B);
    return e;
}

hipError_t Mlp::train_step(const float* X, const float* t, uint32_t B, float lr, float* loss,
                           hipStream_t s) {
    hipError_t e = gradient_dev(X, t, B, loss, s);
    if (e != hipSuccess)
        return e;
    ++step_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)step_), c2 = 1.0f - std::pow(b2, (float)step_);
    hipLaunchKernelGGL(k_adam, dim3(blocks(np_)), dim3(256), 0, s, (float*)params_.p,
                       (const float*)grads_.p, (float*)m_.p, (float*)v_.p, np_, lr, b1, b2, 1e-8f,
                       c1, c2);
    return hipGetLastError();
}

bool Mlp::get_grads(float* host) const {
    return hip_ok(hipMemcpy(host, grads_.p, (size_t)np_ * 4, hipMemcpyDeviceToHost), "mlp grads");
}

} // namespace kfa
