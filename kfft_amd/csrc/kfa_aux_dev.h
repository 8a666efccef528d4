// Per-instance helper operations of the transform plans, as device functions
// shared by their standalone kernels (kfa_aux.hip) and by the tasks of the
// ragged executor (kfa_ragged.hip).  Every function handles one instance b
// (and one bin pair kk for the real rebuild); callers bound b and kk.
#pragma once
#include "kfa_device.h"

namespace kfa {

// out[b] = x_b[0]: the Rader x[0] values, saved so that passes writing the
// output in place never race with readers of x[0]
template <typename T>
__device__ __forceinline__ void gather_first_one(const GatherArgs& a, uint64_t b) {
    const uint64_t hi = b / a.C, lo = b - hi * a.C;
    const int64_t off = (a.tbl ? ((const int64_t*)a.tbl)[hi] : (int64_t)hi * a.dist_hi) +
                        (int64_t)lo * a.dist_lo;
    const T* in = (const T*)a.in;
    T* out = (T*)a.out;
    for (uint32_t c = 0; c < a.comps; ++c)
        out[b * a.comps + c] = in[off * a.comps + c];
}

// Rader DC term: X[0] = scale * (x[0] + conj?(A[0])), A = FFT_{p-1} of the
// gathered input, written through the output layout
template <typename T>
__device__ __forceinline__ void rader_dc_one(const DcArgs& a, uint64_t b) {
    const T* x0 = (const T*)a.x0;
    cx<T> v = ((const cx<T>*)a.A)[b * a.L];
    if (a.conj_)
        v.im = -v.im;
    const cx<T> xv = a.x0_real ? cx<T>{x0[b], T(0)} : cx<T>{x0[2 * b], x0[2 * b + 1]};
    cx<T> d = xv + v;
    d.re *= (T)a.scale;
    d.im *= (T)a.scale;
    const uint64_t hi = b / a.C, lo = b - hi * a.C;
    const int64_t off = (a.tbl ? ((const int64_t*)a.tbl)[hi] : (int64_t)hi * a.dh) +
                        (int64_t)lo * a.dl;
    if (a.real_out)
        ((T*)a.out)[off] = d.re;
    else
        ((cx<T>*)a.out)[off] = d;
}

// Real-input rebuild (even length n = 2h), bins kk and h - kk of instance b
// (kk = 0..h/2): with f1 = Z[k] + conj(Z[h-k]), f2 = Z[k] - conj(Z[h-k]),
// t = f2 * st[k-1]:  X[k] = (f1 + t) / 2,  X[h-k] = conj(f1 - t) / 2; X[0] and
// X[h] from Z[0].  Optional full-spectrum mirror X[n-k] = conj(X[k]).  Safe in
// place (Z == X with the same layout): a call reads and writes only its pair.
template <typename T>
__device__ __forceinline__ void real_post_one(const RealPostArgs& a, uint32_t kk, uint64_t b) {
    const cx<T>* z = (const cx<T>*)a.Z + (a.ztbl ? ((const int64_t*)a.ztbl)[b]
                                                  : (int64_t)b * a.z_dist);
    const uint64_t bh = b / a.C, bl = b - bh * a.C;
    cx<T>* x = (cx<T>*)a.X + (a.xtbl ? ((const int64_t*)a.xtbl)[bh] : (int64_t)bh * a.xdh) +
               (int64_t)bl * a.xdl;
    const uint32_t h = a.h, n = 2 * h;
    const cx<T>* st = (const cx<T>*)a.st;
    auto put = [&](uint32_t k, cx<T> v) {
        if (a.conj_out)
            v.im = -v.im;
        x[(int64_t)k * a.xs] = v;
    };
    if (kk == 0) {
        const cx<T> dc = z[0];
        put(0, {dc.re + dc.im, T(0)});
        put(h, {dc.re - dc.im, T(0)});
        return;
    }
    const uint32_t k = kk;
    cx<T> xk, xh;
    rebuild_pair<T>(z[k], z[h - k], st[k - 1], xk, xh);
    put(k, xk);
    put(h - k, xh);
    if (a.expand) {
        put(n - k, conj(xk));
        put(n - (h - k), conj(xh));
    }
}

} // namespace kfa
