// Stack-based LBVH traversal of one ray by one lane of a wave64, shared by
// the wavefront trace kernel and the path-replay adjoint kernel (which
// re-traces every path vertex and must reproduce the trace stage bit for
// bit).  Semantics follow trace.comp:27-88: opaque geometry, no face
// culling, interval [tmin, tmax] with tmax < tmin a miss, closest hit or any
// hit with early termination.  Closest-hit ties on t resolve to the smaller
// original triangle index, so the result equals a brute-force scan whatever
// the traversal order.
#pragma once

#include "intersect.h"
#include "scene.h"

namespace wfpt {

__device__ __forceinline__ float safe_inv(float d) {
    return 1.0f / (fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d));
}

// Traversal stack: the top LDS_DEPTH entries of each lane live in LDS
// (column per lane, so lanes of a wave hit distinct banks), deeper entries in
// private memory.  The build guarantees depth < TRACE_STACK.
constexpr int LDS_DEPTH = 16;

struct TraceOut {
    int slot;      // leaf slot of the accepted triangle, -1 for a miss
    float t;
    uint32_t nodes_visited, tris_tested;
};

// s_stack: this block's LDS stack (LDS_DEPTH * BLOCK ints) when LDS is set.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS>
__device__ __forceinline__ TraceOut traverse(const float4* __restrict__ nodes,
                                             const float4* __restrict__ tris, float4 ro, float4 rd,
                                             const WtRay& wr, bool live, int* s_stack) {
    TraceOut res{-1, rd.w, 0u, 0u};
    const float tmin = ro.w;
    int best_orig = 0x7fffffff;
    float best_t = rd.w;
    if (live) {
        const float ix = safe_inv(rd.x), iy = safe_inv(rd.y), iz = safe_inv(rd.z);
        const float oix = ro.x * ix, oiy = ro.y * iy, oiz = ro.z * iz;
        int spill[LDS ? TRACE_STACK - LDS_DEPTH : TRACE_STACK];
        int sp = 0;
        const unsigned lane_col = threadIdx.x;
        auto push = [&](int v) {
            if (LDS && sp < LDS_DEPTH) s_stack[sp * BLOCK + lane_col] = v;
            else spill[sp - (LDS ? LDS_DEPTH : 0)] = v;
            ++sp;
        };
        auto pop = [&]() {
            --sp;
            return (LDS && sp < LDS_DEPTH) ? s_stack[sp * BLOCK + lane_col]
                                           : spill[sp - (LDS ? LDS_DEPTH : 0)];
        };
        int node = 0;
        while (true) {
            if (node >= 0) {
                if (STATS) ++res.nodes_visited;
                const float4* nd = nodes + 4 * (int64_t)node;
                const float4 a = nd[0], b = nd[1], c = nd[2], d = nd[3];
                // child 0: lo (a.x a.y a.z) hi (a.w b.x b.y); child 1: lo (b.z b.w c.x) hi (c.y c.z c.w)
                float t0x = fmaf(a.x, ix, -oix), t1x = fmaf(a.w, ix, -oix);
                float t0y = fmaf(a.y, iy, -oiy), t1y = fmaf(b.x, iy, -oiy);
                float t0z = fmaf(a.z, iz, -oiz), t1z = fmaf(b.y, iz, -oiz);
                float n0 = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
                float f0 = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), best_t));
                float u0x = fmaf(b.z, ix, -oix), u1x = fmaf(c.y, ix, -oix);
                float u0y = fmaf(b.w, iy, -oiy), u1y = fmaf(c.z, iy, -oiy);
                float u0z = fmaf(c.x, iz, -oiz), u1z = fmaf(c.w, iz, -oiz);
                float n1 = fmaxf(fmaxf(fminf(u0x, u1x), fminf(u0y, u1y)), fmaxf(fminf(u0z, u1z), tmin));
                float f1 = fminf(fminf(fmaxf(u0x, u1x), fmaxf(u0y, u1y)), fminf(fmaxf(u0z, u1z), best_t));
                const bool h0 = n0 <= f0, h1 = n1 <= f1;
                const int r0 = __float_as_int(d.x), r1 = __float_as_int(d.y);
                if (h0 && h1) {
                    const bool second_first = n1 < n0;
                    node = second_first ? r1 : r0;
                    push(second_first ? r0 : r1);
                } else if (h0) {
                    node = r0;
                } else if (h1) {
                    node = r1;
                } else {
                    if (sp == 0) break;
                    node = pop();
                }
            } else {
                const int first = leaf_first(node), cnt = leaf_count(node);
                bool done = false;
                for (int k = 0; k < cnt; ++k) {
                    const float4* tp = tris + 3 * (int64_t)(first + k);
                    const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                    if (STATS) ++res.tris_tested;
                    WtHit h;
                    if (!wt_intersect(wr, v0, v1, v2, h)) continue;
                    if (!(h.t >= tmin && h.t <= best_t)) continue;
                    const int orig = __float_as_int(v0.w);
                    if (CLOSEST) {
                        if (h.t < best_t || orig < best_orig) {
                            best_t = h.t;
                            best_orig = orig;
                            res.slot = first + k;
                        }
                    } else {
                        best_t = h.t;
                        best_orig = orig;
                        res.slot = first + k;
                        done = true;
                        break;
                    }
                }
                if (done || sp == 0) break;
                node = pop();
            }
        }
    }
    res.t = best_t;
    return res;
}

// Hit record {prim in geometry, geometry, t, packUnorm2x16(bary)} of an
// accepted slot (trace.comp:70-76), or the miss record {-1, -1, -1, 0}
// (trace.comp:78-81).  *orig_tri receives the original triangle index.
__device__ __forceinline__ int4 hit_record(const float4* __restrict__ tris, const WtRay& wr,
                                           const TraceOut& tr, int* orig_tri = nullptr) {
    if (tr.slot < 0) {
        if (orig_tri) *orig_tri = -1;
        return make_int4(-1, -1, __float_as_int(-1.0f), 0);
    }
    const float4* tp = tris + 3 * (int64_t)tr.slot;
    const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
    WtHit h;
    wt_intersect(wr, v0, v1, v2, h);
    if (orig_tri) *orig_tri = __float_as_int(v0.w);
    return make_int4(__float_as_int(v1.w), __float_as_int(v2.w), __float_as_int(tr.t),
                     (int)pack_unorm2x16(h.v / h.det, h.w / h.det));
}

}  // namespace wfpt
