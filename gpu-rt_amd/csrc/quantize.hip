// quantize.hip -- 8-bit quantized copies of the wide nodes.
//
// Each wide node gets an origin (the per-axis minimum of its children's boxes) and per
// axis a power-of-two scale s = 2^(e-127), the smallest for which
// fmaf(255, s, origin) reaches the node's maximum.  A child's box is stored as 8-bit
// multiples of s: lo rounded down, hi rounded up, and both are checked with the exact
// FP32 expression the traversal decodes them with (fmaf(q, s, origin)), so every
// decoded box contains the FP32 box it came from.  FP32 boxes are already inflated for
// rounding, so the culls of the traversal stay conservative and query results do not
// change -- only the node's size does (64 B instead of 128 B at W = 4).  An empty child
// slot is stored as the inverted box lo = 255, hi = 0.
#include <algorithm>

#include "gpurt_internal.h"

namespace gpurt_impl {

__device__ __forceinline__ float pow2_biased(uint32_t e) { return __uint_as_float(e << 23); }

template <int W>
__global__ __launch_bounds__(kBlock) void k_quantize(const WNode<W>* __restrict__ in, QNode<W>* __restrict__ out,
                                                     uint32_t n_host, const uint32_t* __restrict__ n_dev) {
    const uint32_t n = n_dev ? n_dev[0] : n_host;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const WNode<W> nd = in[i];
    float org[3] = {INFINITY, INFINITY, INFINITY};
    float top[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool any = false;
#pragma unroll
    for (int j = 0; j < W; j++) {
        if (nd.child[j] == kEmpty) continue;
        any = true;
        org[0] = fminf(org[0], nd.lox[j]); org[1] = fminf(org[1], nd.loy[j]); org[2] = fminf(org[2], nd.loz[j]);
        top[0] = fmaxf(top[0], nd.hix[j]); top[1] = fmaxf(top[1], nd.hiy[j]); top[2] = fmaxf(top[2], nd.hiz[j]);
    }
    if (!any) {
        org[0] = org[1] = org[2] = 0.0f;
        top[0] = top[1] = top[2] = 0.0f;
    }
    using QN = QNode<W>;
    QN q;
#pragma unroll
    for (int k = 0; k < QN::kWords; k++) q.w[k] = 0u;
    q.w[0] = __float_as_uint(org[0]);
    q.w[1] = __float_as_uint(org[1]);
    q.w[2] = __float_as_uint(org[2]);
    uint32_t exps = 0;
    float sc[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ext = top[a] - org[a];
        // 255 * 2^(k-8) < 2^k: start one binade below the extent's and step up
        int e = (int)((__float_as_uint(ext) >> 23) & 0xFFu) - 8;
        e = e < 1 ? 1 : (e > 254 ? 254 : e);
        while (e < 254 && fmaf(255.0f, pow2_biased((uint32_t)e), org[a]) < top[a]) e++;
        sc[a] = pow2_biased((uint32_t)e);
        exps |= (uint32_t)e << (8 * a);
    }
    if constexpr (QN::kFloatScale) {
        q.w[3] = __float_as_uint(sc[0]);
        q.w[4] = __float_as_uint(sc[1]);
        q.w[5] = __float_as_uint(sc[2]);
    } else {
        q.w[3] = exps;
    }
    uint8_t* const qb = reinterpret_cast<uint8_t*>(q.w + QN::kQ0);  // q[6][W]
#pragma unroll
    for (int j = 0; j < W; j++) {
        q.w[QN::kChild0 + j] = nd.child[j];
        const float lo[3] = {nd.lox[j], nd.loy[j], nd.loz[j]};
        const float hi[3] = {nd.hix[j], nd.hiy[j], nd.hiz[j]};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // an empty slot gets the inverted box [255, 0] on every axis, which the
            // traversal's box test rejects without looking at the child word
            uint32_t ql = 255, qh = 0;
            if (nd.child[j] != kEmpty) {
                // estimates, then exact checks with the decoder's expression; q = 0
                // decodes to origin (<= lo) and q = 255 to >= top (>= hi), so the
                // fall-backs are always valid
                float fl = floorf((lo[a] - org[a]) / sc[a]);
                fl = fminf(fmaxf(fl, 0.0f), 255.0f);
                ql = (uint32_t)fl;
                for (int it = 0; it < 4 && ql > 0 && fmaf((float)ql, sc[a], org[a]) > lo[a]; it++) ql--;
                if (fmaf((float)ql, sc[a], org[a]) > lo[a]) ql = 0;
                float fh = ceilf((hi[a] - org[a]) / sc[a]);
                fh = fminf(fmaxf(fh, 0.0f), 255.0f);
                qh = (uint32_t)fh;
                for (int it = 0; it < 4 && qh < 255 && fmaf((float)qh, sc[a], org[a]) < hi[a]; it++) qh++;
                if (fmaf((float)qh, sc[a], org[a]) < hi[a]) qh = 255;
            }
            qb[a * W + j] = (uint8_t)ql;
            qb[(3 + a) * W + j] = (uint8_t)qh;
        }
    }
    out[i] = q;
}

void launch_quantize(int W, const void* nodes, void* qnodes, uint32_t n_host, uint32_t capacity,
                     const uint32_t* n_dev, hipStream_t s) {
    const uint32_t n = n_dev ? capacity : n_host;
    if (n == 0) return;
    const dim3 grid(cdiv(n, kBlock));
    if (W == 2)
        hipLaunchKernelGGL(k_quantize<2>, grid, dim3(kBlock), 0, s, (const WNode<2>*)nodes, (QNode<2>*)qnodes, n_host, n_dev);
    else if (W == 4)
        hipLaunchKernelGGL(k_quantize<4>, grid, dim3(kBlock), 0, s, (const WNode<4>*)nodes, (QNode<4>*)qnodes, n_host, n_dev);
    else
        hipLaunchKernelGGL(k_quantize<8>, grid, dim3(kBlock), 0, s, (const WNode<8>*)nodes, (QNode<8>*)qnodes, n_host, n_dev);
}

}  // namespace gpurt_impl
