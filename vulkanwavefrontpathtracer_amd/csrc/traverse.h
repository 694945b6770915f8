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

// Per-lane traversal state; the stack top lives in LDS (s_stack column of
// the lane) and the rest in a private spill array kept beside the state
// (SPILL entries).
template <bool LDS, int STACK = TRACE_STACK>
struct TravState {
    static constexpr int SPILL = LDS ? STACK - LDS_DEPTH : STACK;
    float ix, iy, iz, oix, oiy, oiz, tmin, best_t;
    int node, sp, best_slot, best_orig;
    uint32_t nodes_visited, tris_tested;
};

template <bool LDS, int STACK>
__device__ __forceinline__ void trav_begin(TravState<LDS, STACK>& st, float4 ro, float4 rd) {
    st.ix = safe_inv(rd.x);
    st.iy = safe_inv(rd.y);
    st.iz = safe_inv(rd.z);
    st.oix = ro.x * st.ix;
    st.oiy = ro.y * st.iy;
    st.oiz = ro.z * st.iz;
    st.tmin = ro.w;
    st.best_t = rd.w;
    st.node = 0;
    st.sp = 0;
    st.best_slot = -1;
    st.best_orig = 0x7fffffff;
    st.nodes_visited = 0;
    st.tris_tested = 0;
}

// Stack of a lane: top LDS_DEPTH entries in LDS (s_stack column of the
// lane), the rest in the private `spill` array.
template <bool LDS, int BLOCK, int STACK>
__device__ __forceinline__ void stack_push(TravState<LDS, STACK>& st,
                                           int (&spill)[TravState<LDS, STACK>::SPILL], int* s_stack,
                                           int v) {
    if (LDS && st.sp < LDS_DEPTH) s_stack[st.sp * BLOCK + threadIdx.x] = v;
    else spill[st.sp - (LDS ? LDS_DEPTH : 0)] = v;
    ++st.sp;
}

template <bool LDS, int BLOCK, int STACK>
__device__ __forceinline__ int stack_pop(TravState<LDS, STACK>& st,
                                         int (&spill)[TravState<LDS, STACK>::SPILL], int* s_stack) {
    --st.sp;
    int v;
    if (LDS && st.sp < LDS_DEPTH) v = s_stack[st.sp * BLOCK + threadIdx.x];
    else v = spill[st.sp - (LDS ? LDS_DEPTH : 0)];
    return v;
}

// Leaf of up to 8 triangle slots; returns true when an any-hit query is done.
template <bool CLOSEST, bool LDS, int STACK, bool STATS>
__device__ __forceinline__ bool trav_leaf(TravState<LDS, STACK>& st, const float4* __restrict__ tris,
                                          const WtRay& wr) {
    const int first = leaf_first(st.node), cnt = leaf_count(st.node);
    for (int k = 0; k < cnt; ++k) {
        const float4* tp = tris + 3 * (int64_t)(first + k);
        const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
        if (STATS) ++st.tris_tested;
        WtHit h;
        if (!wt_intersect(wr, v0, v1, v2, h)) continue;
        if (!(h.t >= st.tmin && h.t <= st.best_t)) continue;
        const int orig = __float_as_int(v0.w);
        if (CLOSEST) {
            if (h.t < st.best_t || orig < st.best_orig) {
                st.best_t = h.t;
                st.best_orig = orig;
                st.best_slot = first + k;
            }
        } else {
            st.best_t = h.t;
            st.best_orig = orig;
            st.best_slot = first + k;
            return true;
        }
    }
    return false;
}

// One traversal step of the 2-wide hierarchy: an interior node (both child
// boxes) or a whole leaf.  Returns true when the ray is finished (stack
// empty, or any-hit found).
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK>
__device__ __forceinline__ bool trav_step(TravState<LDS, STACK>& st,
                                          int (&spill)[TravState<LDS, STACK>::SPILL],
                                          const float4* __restrict__ nodes,
                                          const float4* __restrict__ tris, const WtRay& wr,
                                          int* s_stack) {
    const float ix = st.ix, iy = st.iy, iz = st.iz, oix = st.oix, oiy = st.oiy, oiz = st.oiz;
    const float tmin = st.tmin;
    if (st.node >= 0) {
        if (STATS) ++st.nodes_visited;
        const float4* nd = nodes + 4 * (int64_t)st.node;
        const float4 a = nd[0], b = nd[1], c = nd[2], d = nd[3];
        const float best_t = st.best_t;
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
            st.node = second_first ? r1 : r0;
            stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, second_first ? r0 : r1);
        } else if (h0) {
            st.node = r0;
        } else if (h1) {
            st.node = r1;
        } else {
            if (st.sp == 0) return true;
            st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
        }
        return false;
    }
    if (trav_leaf<CLOSEST, LDS, STACK, STATS>(st, tris, wr)) return true;
    if (st.sp == 0) return true;
    st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    return false;
}

__device__ __forceinline__ void cswap(float& ta, int& ra, float& tb, int& rb) {
    if (tb < ta) {
        const float t = ta;
        ta = tb;
        tb = t;
        const int r = ra;
        ra = rb;
        rb = r;
    }
}

// One traversal step of the 4-wide quantized hierarchy: up to four child
// boxes decoded from 8-bit planes (t = q * 2^e / d + (origin - o) / d, one
// fused multiply-add per plane), the hit children visited nearest first.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK>
__device__ __forceinline__ bool trav_step4(TravState<LDS, STACK>& st,
                                           int (&spill)[TravState<LDS, STACK>::SPILL],
                                           const float4* __restrict__ nodes4,
                                           const float4* __restrict__ tris, const WtRay& wr,
                                           int* s_stack) {
    if (st.node >= 0) {
        if (STATS) ++st.nodes_visited;
        const float4* nd = nodes4 + 4 * (int64_t)st.node;
        const float4 a = nd[0], b = nd[1], c = nd[2], d = nd[3];
        const uint32_t meta = __float_as_uint(a.w);
        const int nc = (int)(meta >> 24);
        const float sax = __uint_as_float((meta & 0xffu) << 23) * st.ix;
        const float say = __uint_as_float(((meta >> 8) & 0xffu) << 23) * st.iy;
        const float saz = __uint_as_float(((meta >> 16) & 0xffu) << 23) * st.iz;
        const float bx = fmaf(a.x, st.ix, -st.oix), by = fmaf(a.y, st.iy, -st.oiy),
                    bz = fmaf(a.z, st.iz, -st.oiz);
        const uint32_t qlx = __float_as_uint(b.x), qly = __float_as_uint(b.y),
                       qlz = __float_as_uint(b.z), qhx = __float_as_uint(b.w),
                       qhy = __float_as_uint(c.x), qhz = __float_as_uint(c.y);
        float tn[4];
        int ref[4] = {__float_as_int(c.z), __float_as_int(c.w), __float_as_int(d.x), __float_as_int(d.y)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int sh = 8 * k;
            const float t0x = fmaf((float)((qlx >> sh) & 0xffu), sax, bx);
            const float t1x = fmaf((float)((qhx >> sh) & 0xffu), sax, bx);
            const float t0y = fmaf((float)((qly >> sh) & 0xffu), say, by);
            const float t1y = fmaf((float)((qhy >> sh) & 0xffu), say, by);
            const float t0z = fmaf((float)((qlz >> sh) & 0xffu), saz, bz);
            const float t1z = fmaf((float)((qhz >> sh) & 0xffu), saz, bz);
            const float nn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), st.tmin));
            const float ff = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), st.best_t));
            tn[k] = (k < nc && nn <= ff) ? nn : INFINITY;
        }
        cswap(tn[0], ref[0], tn[1], ref[1]);
        cswap(tn[2], ref[2], tn[3], ref[3]);
        cswap(tn[0], ref[0], tn[2], ref[2]);
        cswap(tn[1], ref[1], tn[3], ref[3]);
        cswap(tn[1], ref[1], tn[2], ref[2]);
        if (tn[3] < INFINITY) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[3]);
        if (tn[2] < INFINITY) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[2]);
        if (tn[1] < INFINITY) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[1]);
        if (tn[0] < INFINITY) {
            st.node = ref[0];
        } else {
            if (st.sp == 0) return true;
            st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
        }
        return false;
    }
    if (trav_leaf<CLOSEST, LDS, STACK, STATS>(st, tris, wr)) return true;
    if (st.sp == 0) return true;
    st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    return false;
}

// s_stack: this block's LDS stack (LDS_DEPTH * BLOCK ints) when LDS is set.
// WIDE selects the 4-wide quantized hierarchy (nodes = scene nodes4).
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, bool WIDE = false>
__device__ __forceinline__ TraceOut traverse(const float4* __restrict__ nodes,
                                             const float4* __restrict__ tris, float4 ro, float4 rd,
                                             const WtRay& wr, bool live, int* s_stack) {
    constexpr int STACK = WIDE ? TRACE_STACK4 : TRACE_STACK;
    TravState<LDS, STACK> st;
    int spill[TravState<LDS, STACK>::SPILL];
    trav_begin(st, ro, rd);
    if (live) {
        if (WIDE) {
            while (!trav_step4<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack)) {
            }
        } else {
            while (!trav_step<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack)) {
            }
        }
    }
    return TraceOut{st.best_slot, st.best_t, st.nodes_visited, st.tris_tested};
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
