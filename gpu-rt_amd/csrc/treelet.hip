// treelet.hip -- SAH treelet restructuring of the binary LBVH (Karras & Aila, "Fast
// Parallel Construction of High-Quality Bounding Volume Hierarchies", HPG 2013).
//
// For every internal node r with more than 2 leaf_size triangles, bottom-up:
//   1. form a treelet: start from r's two children and repeatedly open the treelet leaf
//      (an internal node) with the largest surface area, up to N leaves (N = 5..7), one
//      candidate per lane;
//   2. find the binary topology over those leaves with the least SAH cost by dynamic
//      programming over all 2^N - 1 leaf subsets (one 64-lane wave per treelet: lanes
//      own subsets; each subset tries every split that keeps its lowest leaf on one
//      side);
//   3. if it is cheaper than the current one, rewrite the treelet's N - 1 internal nodes
//      (r keeps its index, so ancestors are untouched) with their new children, boxes,
//      triangle counts and costs.
// Bottom-up order is the same as the box pass: one launch per Karras prefix length,
// largest first.  Two roots of one launch have equal prefix length, so neither is an
// ancestor of the other and their treelets are disjoint; everything below a root was
// finished by earlier launches.  (That argument holds for the tree as Karras built it,
// so one round is supported.)  Subtrees of <= leaf_size triangles become single leaves
// in the wide collapse and are priced as leaves, so they are never restructured.
#include <algorithm>

#include "build_common.h"

namespace gpurt_impl {

// the s-th submask of mask in increasing order (bit i of s selects the i-th set bit)
__device__ __forceinline__ int nth_submask(int s, int mask) {
    int out = 0;
    for (int bit = 0; mask; bit++) {
        const int lo = mask & -mask;
        if (s & (1 << bit)) out |= lo;
        mask ^= lo;
    }
    return out;
}

template <int N>
__global__ __launch_bounds__(64) void k_treelet(int level, const uint32_t* __restrict__ lvl_begin,
                                                const uint32_t* __restrict__ lvl_order, int leaf_size,
                                                uint32_t* left, uint32_t* right, uint32_t* cnt, float* cost,
                                                float4* nbox, const float4* __restrict__ saabb) {
    constexpr int NS = 1 << N;
    __shared__ uint32_t s_leaf[N], s_int[N];
    __shared__ float4 s_llo[N], s_lhi[N];
    __shared__ uint32_t s_lcnt[N];
    __shared__ float s_lcost[N];
    __shared__ float4 s_lo[NS], s_hi[NS];
    __shared__ uint32_t s_cnt[NS];
    __shared__ float s_copt[NS];
    __shared__ uint8_t s_part[NS];
    __shared__ uint8_t s_list[NS];  // subsets ordered by size: size k at [s_kbase[k], + C(nl, k))
    __shared__ int s_kbase[N + 2], s_kfill[N + 2];
    __shared__ float s_cur;
    const int lane = threadIdx.x;
    // one wave per treelet root of this prefix length: the roots with more than
    // 2 leaf_size triangles (bucket 2 * level; the others become at most two leaves of one
    // wide node, where a new topology gains nothing measurable -- rays_scene: skipping them
    // builds in 10.7 instead of 13.4 ms with the same ray times; skipping up to 16 or 32
    // triangles builds faster still but costs 2-5% in ray time)
    const uint32_t beg = lvl_begin[2 * level], end = lvl_begin[2 * level + 1];
    for (uint32_t k = beg + blockIdx.x; k < end; k += gridDim.x) {
        const int r = (int)lvl_order[k];
        {
            // Form the treelet with one candidate (treelet leaf) per lane: a lane that
            // takes a candidate loads its box, triangle count, cost and, for an internal
            // node, its children, all at once, so opening a node costs one round of
            // loads (by the two lanes that receive its children) instead of a serial
            // scan of every candidate's box.  The node opened is the internal candidate
            // with the largest surface area, the lowest slot on ties.
            uint32_t x = kEmpty, xl = 0, xr = 0, xc = 0;
            float4 xlo = make_float4(0.f, 0.f, 0.f, 0.f), xhi = xlo;
            float xcost = 0.0f, xarea = -1.0f;
            auto take = [&](uint32_t ref) {
                x = ref;
                ref_box(ref, saabb, nbox, xlo, xhi);
                if (ref & kLeafBit) {
                    xc = 1u;
                    xcost = kCt * box_area(xlo, xhi);
                    xarea = -1.0f;  // not openable
                } else {
                    xc = cnt[ref];
                    xcost = cost[ref];
                    xl = left[ref];
                    xr = right[ref];
                    const float a = box_area(xlo, xhi);
                    xarea = a > -1.0f ? a : -1.0f;  // (a NaN area is never opened)
                }
            };
            float4 rlo = xlo, rhi = xhi;
            uint32_t rc = 0;
            if (lane == 0) {
                rlo = nbox[2 * (size_t)r];
                rhi = nbox[2 * (size_t)r + 1];
                rc = cnt[r];
            }
            if (lane < 2) take(lane ? right[r] : left[r]);
            // the root's cost, refreshed from its (finished) children, the first two
            // candidates
            const float c_right = __shfl(xcost, 1);
            if (lane == 0) {
                s_cur = sah_node_cost(rlo, rhi, rc, xcost, c_right, leaf_size);
                cost[r] = s_cur;
                s_int[0] = (uint32_t)r;
            }
            int nl = 2, ni = 1;
            while (nl < N) {
                float a = lane < nl ? xarea : -1.0f;
                int j = lane;
#pragma unroll
                for (int o = 4; o >= 1; o >>= 1) {  // N <= 7: lanes 0..7 hold every candidate
                    const float oa = __shfl_xor(a, o);
                    const int oj = __shfl_xor(j, o);
                    if (oa > a || (oa == a && oj < j)) {
                        a = oa;
                        j = oj;
                    }
                }
                if (__shfl(a, 0) < 0.0f) break;  // no internal candidate left
                const int jb = __shfl(j, 0);
                const uint32_t ob = __shfl(x, jb), lb = __shfl(xl, jb), rb = __shfl(xr, jb);
                if (lane == 0) s_int[ni] = ob;
                ni++;
                if (lane == jb) take(lb);
                if (lane == nl) take(rb);
                nl++;
            }
            // every lane ran the same loop: nl is wave-uniform, so the table lookups below
            // are scalar loads rather than one dependent vector load per lane
            nl = __builtin_amdgcn_readfirstlane(nl);
            if (lane < nl) {
                s_leaf[lane] = x;
                s_llo[lane] = xlo;
                s_lhi[lane] = xhi;
                s_lcnt[lane] = xc;
                s_lcost[lane] = xcost;
            }
            __syncthreads();
            const int full = (1 << nl) - 1;
            for (int sub = lane + 1; sub <= full; sub += 64) {
                float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
                float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
                uint32_t c = 0;
                for (int j = 0; j < nl; j++) {
                    if (!(sub & (1 << j))) continue;
                    lo.x = fminf(lo.x, s_llo[j].x); lo.y = fminf(lo.y, s_llo[j].y); lo.z = fminf(lo.z, s_llo[j].z);
                    hi.x = fmaxf(hi.x, s_lhi[j].x); hi.y = fmaxf(hi.y, s_lhi[j].y); hi.z = fmaxf(hi.z, s_lhi[j].z);
                    c += s_lcnt[j];
                }
                s_lo[sub] = lo;
                s_hi[sub] = hi;
                s_cnt[sub] = c;
                if ((sub & (sub - 1)) == 0) {
                    s_copt[sub] = s_lcost[__ffs(sub) - 1];
                    s_part[sub] = 0;
                }
            }
            if (lane <= nl) {
                // subsets of fewer than `lane` leaves (of at least one): the sum of C(nl, j)
                // for 1 <= j < lane, with C(nl, j) = C(nl, j - 1) (nl - j + 1) / j
                int b = 0, c = 1;
                for (int j = 1; j < lane; j++) {
                    c = c * (nl - j + 1) / j;
                    b += c;
                }
                s_kbase[lane] = b;
                s_kfill[lane] = 0;
            }
            __syncthreads();
            for (int sub = lane + 1; sub <= full; sub += 64) {
                const int k = __popc(sub);
                s_list[s_kbase[k] + atomicAdd(&s_kfill[k], 1)] = (uint8_t)sub;
            }
            __syncthreads();
            // Optimal costs by subset size.  A subset's splits keep its lowest leaf on one
            // side (P holds it, P != sub).  Up to 4 leaves a lane takes a subset and tries
            // its <= 7 splits; from 5 leaves a group of 2^(k-1) lanes takes one subset,
            // a lane per split, and reduces.  Ties go to the larger P in both forms (the
            // first found in a descending walk of the submasks), so the tree does not
            // depend on which form ran.
            int nk = nl;  // C(nl, k), updated at the top of each step
            for (int k = 2; k <= nl; k++) {
                nk = nk * (nl - k + 1) / k;
                const uint8_t* list = s_list + s_kbase[k];
                if (k <= 4) {
                    for (int idx = lane; idx < nk; idx += 64) {
                        const int sub = list[idx];
                        const int low = sub & -sub;
                        float best = INFINITY;
                        int bp = low;
                        for (int P = (sub - 1) & sub; P; P = (P - 1) & sub) {
                            if (!(P & low)) continue;
                            const float c = s_copt[P] + s_copt[sub ^ P];
                            if (c < best) {
                                best = c;
                                bp = P;
                            }
                        }
                        const float a = box_area(s_lo[sub], s_hi[sub]);
                        s_copt[sub] =
                            (s_cnt[sub] <= (uint32_t)leaf_size) ? kCt * a * (float)s_cnt[sub] : kCi * a + best;
                        s_part[sub] = (uint8_t)bp;
                    }
                } else {
                    const int G = 1 << (k - 1);  // 16, 32 or 64 lanes per subset
                    const int groups = 64 / G, g = lane / G, sl = lane % G;
                    for (int base = 0; base < nk; base += groups) {
                        const int idx = base + g;
                        const bool ok = idx < nk;
                        const int sub = ok ? list[idx] : 0;
                        float c = INFINITY;
                        int P = -1;
                        if (ok && sl < G - 1) {
                            const int low = sub & -sub;
                            P = low | nth_submask(sl, sub ^ low);
                            c = s_copt[P] + s_copt[sub ^ P];
                        }
                        for (int o = G / 2; o >= 1; o >>= 1) {
                            const float oc = __shfl_xor(c, o);
                            const int op = __shfl_xor(P, o);
                            if (oc < c || (oc == c && op > P)) {
                                c = oc;
                                P = op;
                            }
                        }
                        if (ok && sl == 0) {
                            const float a = box_area(s_lo[sub], s_hi[sub]);
                            s_copt[sub] =
                                (s_cnt[sub] <= (uint32_t)leaf_size) ? kCt * a * (float)s_cnt[sub] : kCi * a + c;
                            s_part[sub] = (uint8_t)P;
                        }
                    }
                }
                __syncthreads();
            }
            if (lane == 0 && nl >= 3 && s_copt[full] < s_cur * (1.0f - 1e-5f)) {
                // rebuild the treelet top-down from the optimal partitions, reusing its
                // internal node indices (r first)
                int st_sub[N], st_node[N];
                int sp = 0, pool = 1;
                st_sub[sp] = full;
                st_node[sp] = r;
                sp++;
                while (sp > 0) {
                    sp--;
                    const int sub = st_sub[sp];
                    const int node = st_node[sp];
                    const int P = s_part[sub];
                    const int Q = sub ^ P;
                    uint32_t ref[2];
                    const int parts[2] = {P, Q};
                    for (int h = 0; h < 2; h++) {
                        const int part = parts[h];
                        if ((part & (part - 1)) == 0) {
                            ref[h] = s_leaf[__ffs(part) - 1];
                        } else {
                            const uint32_t idx = s_int[pool++];
                            ref[h] = idx;
                            st_sub[sp] = part;
                            st_node[sp] = (int)idx;
                            sp++;
                        }
                    }
                    left[node] = ref[0];
                    right[node] = ref[1];
                    nbox[2 * (size_t)node] = s_lo[sub];
                    nbox[2 * (size_t)node + 1] = s_hi[sub];
                    cnt[node] = s_cnt[sub];
                    cost[node] = s_copt[sub];
                }
            }
            __syncthreads();
        }
    }
}

template <int N>
static void launch_treelet_levels(BuildBufs& b, int key_bits, int nbits, hipStream_t s) {
    const unsigned grid = std::min(cdiv(b.N, 64), 8192u);
    for (int level = 95; level >= 64 - key_bits; level--) {
        if (level >= 64 && level < 96 - nbits) continue;
        hipLaunchKernelGGL(k_treelet<N>, dim3(grid), dim3(64), 0, s, level, b.lvl_begin, b.lvl_order, b.leaf_size,
                           b.left, b.right, b.cnt, b.cost, b.nbox, b.saabb);
    }
}

int run_treelets(BuildBufs& b, int treelet_size, int rounds, int key_bits, int nbits, hipStream_t s) {
    // buckets (b.lvl_begin / b.lvl_order) were filled by the box pass
    if (rounds > 1)
        return set_error(GPURT_E_UNSUPPORTED, "treelet_rounds > 1 is not supported (one bottom-up round)");
    if (b.N < 3) return GPURT_OK;
    switch (treelet_size) {
        case 5: launch_treelet_levels<5>(b, key_bits, nbits, s); break;
        case 6: launch_treelet_levels<6>(b, key_bits, nbits, s); break;
        case 7: launch_treelet_levels<7>(b, key_bits, nbits, s); break;
        default: return set_error(GPURT_E_UNSUPPORTED, "treelet_size must be 5..7");
    }
    GPURT_HIP(hipGetLastError());
    return GPURT_OK;
}

}  // namespace gpurt_impl
