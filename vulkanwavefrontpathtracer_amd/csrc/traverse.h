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
typedef __attribute__((address_space(3))) int lds_int;

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
    // hit record fields kept at acceptance (RECORD traversals only)
    int best_prim, best_geom;
    uint32_t best_bary;
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
    st.best_prim = -1;
    st.best_geom = -1;
    st.best_bary = 0u;
}

// Stack of a lane: top LDS_DEPTH entries in LDS (s_stack column of the
// lane), the rest in the private `spill` array.
template <bool LDS, int BLOCK, int STACK>
__device__ __forceinline__ void stack_push(TravState<LDS, STACK>& st,
                                           int (&spill)[TravState<LDS, STACK>::SPILL], int* s_stack,
                                           int v) {
    if (LDS && st.sp < LDS_DEPTH) ((lds_int*)s_stack)[st.sp * BLOCK + threadIdx.x] = v;
    else spill[st.sp - (LDS ? LDS_DEPTH : 0)] = v;
    ++st.sp;
}

// LDS entries are read through an address-space-3 pointer: through the
// generic pointer the compiler merges the LDS and the private load of the
// two branches into one flat load of a selected address, which goes through
// the vector memory pipeline (and its wait counter) even for LDS entries.
template <bool LDS, int BLOCK, int STACK>
__device__ __forceinline__ int stack_pop(TravState<LDS, STACK>& st,
                                         int (&spill)[TravState<LDS, STACK>::SPILL], int* s_stack) {
    --st.sp;
    if (!LDS) return spill[st.sp];
    int v = ((lds_int*)s_stack)[min(st.sp, LDS_DEPTH - 1) * BLOCK + threadIdx.x];
    if (st.sp >= LDS_DEPTH) v = spill[st.sp - LDS_DEPTH];
    return v;
}

// Triangle in leaf slot `slot` (vertices v0..v2, original index in v0.w);
// returns true when an any-hit query is done.  RECORD also keeps the hit
// record fields of the accepted triangle (prim in v1.w, geometry in v2.w,
// packed barycentrics as hit_record computes them), so a finished lane can
// store its record without reloading the triangle.
template <bool CLOSEST, bool LDS, int STACK, bool STATS, bool RECORD = false>
__device__ __forceinline__ bool trav_tri(TravState<LDS, STACK>& st, int slot, float4 v0, float4 v1,
                                         float4 v2, const WtRay& wr) {
    if (STATS) ++st.tris_tested;
    WtHit h;
    if (!wt_intersect(wr, v0, v1, v2, h)) return false;
    if (!(h.t >= st.tmin && h.t <= st.best_t)) return false;
    const int orig = __float_as_int(v0.w);
    if (CLOSEST && !(h.t < st.best_t || orig < st.best_orig)) return false;
    st.best_t = h.t;
    st.best_orig = orig;
    st.best_slot = slot;
    if (RECORD) {
        st.best_prim = __float_as_int(v1.w);
        st.best_geom = __float_as_int(v2.w);
        st.best_bary = pack_unorm2x16(h.v / h.det, h.w / h.det);
    }
    return !CLOSEST;
}

// Leaf `leaf` (reference of up to 8 triangle slots); returns true when an
// any-hit query is done.
template <bool CLOSEST, bool LDS, int STACK, bool STATS, bool RECORD = false>
__device__ __forceinline__ bool trav_leaf(TravState<LDS, STACK>& st, int leaf,
                                          const float4* __restrict__ tris, const WtRay& wr) {
    const int first = leaf_first(leaf), cnt = leaf_count(leaf);
    for (int k = 0; k < cnt; ++k) {
        const float4* tp = tris + 3 * (int64_t)(first + k);
        if (trav_tri<CLOSEST, LDS, STACK, STATS, RECORD>(st, first + k, tp[0], tp[1], tp[2], wr))
            return true;
    }
    return false;
}

// One traversal step of the 2-wide hierarchy: an interior node (both child
// boxes) or a whole leaf.  Returns true when the ray is finished (stack
// empty, or any-hit found).
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK, bool RECORD = false>
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
    if (trav_leaf<CLOSEST, LDS, STACK, STATS, RECORD>(st, st.node, tris, wr)) return true;
    if (st.sp == 0) return true;
    st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    return false;
}

// Compare-exchange of (distance, reference) pairs, written as selects so the
// sorting network stays branch-free (one compare and four selects).
__device__ __forceinline__ void cswap(float& ta, int& ra, float& tb, int& rb) {
    const bool c = tb < ta;
    const float lo = c ? tb : ta, hi = c ? ta : tb;
    const int rl = c ? rb : ra, rh = c ? ra : rb;
    ta = lo;
    tb = hi;
    ra = rl;
    rb = rh;
}

// Interior node st.node of the 4-wide quantized hierarchy: up to four child
// boxes decoded from 8-bit planes (t = q * 2^e / d + (origin - o) / d, one
// fused multiply-add per plane), the hit children visited nearest first.
// Returns false when the lane has no node left (no child hit, stack empty).
// Child boxes of a 4-wide node (a..d) hit within [tmin, best_t], nearest
// first: tn[k] the entry distance (INFINITY for a miss), ref[k] the child.
template <bool LDS, int STACK>
__device__ __forceinline__ void node4_hits(const TravState<LDS, STACK>& st, float4 a, float4 b,
                                           float4 c, float4 d, float (&tn)[4], int (&ref)[4]) {
    // empty child slots hold inverted boxes and a harmless reference (see
    // wide_emit_kernel), so every slot is tested without the child count
    const uint32_t meta = __float_as_uint(a.w);
    const float sax = __uint_as_float((meta & 0xffu) << 23) * st.ix;
    const float say = __uint_as_float(((meta >> 8) & 0xffu) << 23) * st.iy;
    const float saz = __uint_as_float(((meta >> 16) & 0xffu) << 23) * st.iz;
    const float bx = fmaf(a.x, st.ix, -st.oix), by = fmaf(a.y, st.iy, -st.oiy),
                bz = fmaf(a.z, st.iz, -st.oiz);
    // Near and far planes per axis by the sign of the inverse direction (a
    // positive one meets the low plane first): the same floats as the min
    // and max of the two plane distances, chosen once per node instead of
    // per child.
    const uint32_t qlx = __float_as_uint(b.x), qly = __float_as_uint(b.y),
                   qlz = __float_as_uint(b.z), qhx = __float_as_uint(b.w),
                   qhy = __float_as_uint(c.x), qhz = __float_as_uint(c.y);
    const bool fx = st.ix < 0.0f, fy = st.iy < 0.0f, fz = st.iz < 0.0f;
    const uint32_t qnx = fx ? qhx : qlx, qfx = fx ? qlx : qhx;
    const uint32_t qny = fy ? qhy : qly, qfy = fy ? qly : qhy;
    const uint32_t qnz = fz ? qhz : qlz, qfz = fz ? qlz : qhz;
    ref[0] = __float_as_int(c.z);
    ref[1] = __float_as_int(c.w);
    ref[2] = __float_as_int(d.x);
    ref[3] = __float_as_int(d.y);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int sh = 8 * k;
        const float nx = fmaf((float)((qnx >> sh) & 0xffu), sax, bx);
        const float fxv = fmaf((float)((qfx >> sh) & 0xffu), sax, bx);
        const float ny = fmaf((float)((qny >> sh) & 0xffu), say, by);
        const float fyv = fmaf((float)((qfy >> sh) & 0xffu), say, by);
        const float nz = fmaf((float)((qnz >> sh) & 0xffu), saz, bz);
        const float fzv = fmaf((float)((qfz >> sh) & 0xffu), saz, bz);
        const float nn = fmaxf(fmaxf(nx, ny), fmaxf(nz, st.tmin));
        const float ff = fminf(fminf(fxv, fyv), fminf(fzv, st.best_t));
        tn[k] = nn <= ff ? nn : INFINITY;
    }
    cswap(tn[0], ref[0], tn[1], ref[1]);
    cswap(tn[2], ref[2], tn[3], ref[3]);
    cswap(tn[0], ref[0], tn[2], ref[2]);
    cswap(tn[1], ref[1], tn[3], ref[3]);
    cswap(tn[1], ref[1], tn[2], ref[2]);
}

// TOP: `top` holds the stack's top entry read before this node's pushes (the
// fused step reads it beside the node fetch); a node that pushes nothing and
// pops takes it instead of reading LDS again.
template <int BLOCK, bool LDS, bool STATS, int STACK, bool TOP = false>
__device__ __forceinline__ bool trav_node4_data(TravState<LDS, STACK>& st,
                                                int (&spill)[TravState<LDS, STACK>::SPILL],
                                                float4 a, float4 b, float4 c, float4 d, int* s_stack,
                                                int top = 0) {
    if (STATS) ++st.nodes_visited;
    float tn[4];
    int ref[4];
    node4_hits(st, a, b, c, d, tn, ref);
    const int h1 = tn[1] < INFINITY, h2 = tn[2] < INFINITY, h3 = tn[3] < INFINITY;
    if (LDS && __all(st.sp <= LDS_DEPTH - 3)) {
        // The hit children come first (nearest first), so the pushes can be
        // three unconditional LDS writes: a missed child's slot is
        // overwritten by the next write or left above the new top.
        lds_int* col = (lds_int*)s_stack + threadIdx.x;
        col[st.sp * BLOCK] = ref[3];
        col[(st.sp + h3) * BLOCK] = ref[2];
        col[(st.sp + h3 + h2) * BLOCK] = ref[1];
        st.sp += h1 + h2 + h3;
    } else {
        if (h3) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[3]);
        if (h2) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[2]);
        if (h1) stack_push<LDS, BLOCK, STACK>(st, spill, s_stack, ref[1]);
    }
    if (tn[0] < INFINITY) {
        st.node = ref[0];
        return true;
    }
    // no child hit, so nothing was pushed and `top` is still the top entry
    if (st.sp == 0) return false;
    if (TOP && LDS && st.sp <= LDS_DEPTH) {
        --st.sp;
        st.node = top;
    } else {
        st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    }
    return true;
}

template <int BLOCK, bool LDS, bool STATS, int STACK>
__device__ __forceinline__ bool trav_node4(TravState<LDS, STACK>& st,
                                           int (&spill)[TravState<LDS, STACK>::SPILL],
                                           const float4* __restrict__ nodes4, int* s_stack) {
    const float4* nd = nodes4 + 4 * (int64_t)st.node;
    return trav_node4_data<BLOCK, LDS, STATS>(st, spill, nd[0], nd[1], nd[2], nd[3], s_stack);
}

// One traversal step of the 4-wide hierarchy that issues a single load per
// lane whatever the lane visits: the node's 64 bytes, or the first triangle
// of a leaf (48 bytes), so a wave whose lanes are split between nodes and
// leaves waits for memory once per step instead of once per branch.
// Returns true when the ray is finished.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK, bool RECORD = false>
__device__ __forceinline__ bool trav_step4_fused(TravState<LDS, STACK>& st,
                                                 int (&spill)[TravState<LDS, STACK>::SPILL],
                                                 const float4* __restrict__ nodes4,
                                                 const float4* __restrict__ tris, const WtRay& wr,
                                                 int* s_stack) {
    const bool inner = st.node >= 0;
    const float4* p = inner ? nodes4 + 4 * (int64_t)st.node : tris + 3 * (int64_t)leaf_first(st.node);
    const float4 a = p[0], b = p[1], c = p[2];
    float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
    if (inner) d = p[3];
    // The stack's top entry, read while the fetch is in flight: the next node
    // of a step that pushes nothing (a leaf, or a node with no child hit),
    // which then waits for neither LDS nor the fetch's dependents to pop.
    int top = 0;
    if (LDS && st.sp > 0) top = ((const lds_int*)s_stack)[min(st.sp - 1, LDS_DEPTH - 1) * BLOCK + threadIdx.x];
    if (inner) return !trav_node4_data<BLOCK, LDS, STATS, STACK, true>(st, spill, a, b, c, d, s_stack, top);
    const int first = leaf_first(st.node), cnt = leaf_count(st.node);
    if (trav_tri<CLOSEST, LDS, STACK, STATS, RECORD>(st, first, a, b, c, wr)) return true;
    for (int k = 1; k < cnt; ++k) {
        const float4* tp = tris + 3 * (int64_t)(first + k);
        if (trav_tri<CLOSEST, LDS, STACK, STATS, RECORD>(st, first + k, tp[0], tp[1], tp[2], wr))
            return true;
    }
    if (st.sp == 0) return true;
    if (LDS && st.sp <= LDS_DEPTH) {
        --st.sp;
        st.node = top;
    } else {
        st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    }
    return false;
}

// One traversal step of the 4-wide hierarchy: an interior node or a whole
// leaf.  Returns true when the ray is finished (stack empty, or any-hit
// found).
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK, bool RECORD = false>
__device__ __forceinline__ bool trav_step4(TravState<LDS, STACK>& st,
                                           int (&spill)[TravState<LDS, STACK>::SPILL],
                                           const float4* __restrict__ nodes4,
                                           const float4* __restrict__ tris, const WtRay& wr,
                                           int* s_stack) {
    if (st.node >= 0) return !trav_node4<BLOCK, LDS, STATS>(st, spill, nodes4, s_stack);
    if (trav_leaf<CLOSEST, LDS, STACK, STATS, RECORD>(st, st.node, tris, wr)) return true;
    if (st.sp == 0) return true;
    st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
    return false;
}

// 4-wide traversal with postponed leaves (Aila & Laine, "Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009, speculative traversal): a
// lane that reaches a leaf parks it and keeps descending interior nodes; the
// wave tests triangles only once every lane still descending has a parked
// leaf, so the triangle test is not issued for one or two lanes on most
// steps.  Order-independent result: the closest hit (ties to the smaller
// original triangle index) or any hit.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK>
__device__ __forceinline__ void trav4_postponed(TravState<LDS, STACK>& st,
                                                int (&spill)[TravState<LDS, STACK>::SPILL],
                                                const float4* __restrict__ nodes4,
                                                const float4* __restrict__ tris, const WtRay& wr,
                                                int* s_stack, bool live) {
    bool have = live;  // st.node holds a node (interior or leaf) still to visit
    int leaf = 0;      // parked leaf reference (< 0), 0 for none
    while (have || leaf != 0) {
        while (have && (st.node >= 0 || leaf == 0)) {
            if (st.node < 0) {
                leaf = st.node;
                have = st.sp > 0;
                if (have) st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
            } else {
                have = trav_node4<BLOCK, LDS, STATS>(st, spill, nodes4, s_stack);
            }
            if (__all(!have || leaf != 0)) break;
        }
        while (leaf != 0) {
            if (trav_leaf<CLOSEST, LDS, STACK, STATS>(st, leaf, tris, wr)) {
                have = false;
                leaf = 0;
                break;
            }
            if (have && st.node < 0) {
                leaf = st.node;
                have = st.sp > 0;
                if (have) st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
            } else {
                leaf = 0;
            }
        }
    }
}

// 4-wide traversal with gated leaves: a lane whose next entry is a leaf
// waits (it neither tests nor moves) while the wave steps its interior
// nodes, and the wave runs the triangle test only once `gate` or more lanes
// wait, or when no lane has an interior node to step.  With one-triangle
// leaves some lane sits at a leaf on most steps, so without the gate the
// triangle test is issued nearly every step for one or two lanes; a waiting
// lane keeps its place, so no traversal is done before its closest hit is
// known (unlike postponed leaves).  gate = 1 is the plain loop.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, int STACK>
__device__ __forceinline__ void trav4_gated(TravState<LDS, STACK>& st,
                                            int (&spill)[TravState<LDS, STACK>::SPILL],
                                            const float4* __restrict__ nodes4,
                                            const float4* __restrict__ tris, const WtRay& wr,
                                            int* s_stack, bool live, int gate) {
    bool active = live;
    while (true) {
        const bool inner = active && st.node >= 0;
        const uint64_t at_inner = ballot64(inner);
        const uint64_t at_leaf = ballot64(active && st.node < 0);
        if ((at_inner | at_leaf) == 0) break;
        if (inner) active = trav_node4<BLOCK, LDS, STATS>(st, spill, nodes4, s_stack);
        if (at_leaf != 0 && (__popcll(at_leaf) >= gate || at_inner == 0)) {
            if (active && st.node < 0 && ((at_leaf >> lane_id()) & 1ull)) {
                if (trav_leaf<CLOSEST, LDS, STACK, STATS>(st, st.node, tris, wr)) {
                    active = false;
                } else if (st.sp == 0) {
                    active = false;
                } else {
                    st.node = stack_pop<LDS, BLOCK, STACK>(st, spill, s_stack);
                }
            }
        }
    }
}

// Loop structures of the 4-wide traversal (the trace.loop knob).
enum TraceLoop { LOOP_BRANCH = 0, LOOP_FUSED = 1, LOOP_POSTPONED = 2, LOOP_GATED = 3 };
constexpr int TRACE_LOOP_DEFAULT = LOOP_FUSED;

// s_stack: this block's LDS stack (LDS_DEPTH * BLOCK ints) when LDS is set.
// WIDE selects the 4-wide quantized hierarchy (nodes = scene nodes4) and
// LOOP its loop structure.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS, bool WIDE = false,
          int LOOP = TRACE_LOOP_DEFAULT>
__device__ __forceinline__ TraceOut traverse(const float4* __restrict__ nodes,
                                             const float4* __restrict__ tris, float4 ro, float4 rd,
                                             const WtRay& wr, bool live, int* s_stack,
                                             int leaf_gate = 1) {
    constexpr int STACK = WIDE ? TRACE_STACK4 : TRACE_STACK;
    TravState<LDS, STACK> st;
    int spill[TravState<LDS, STACK>::SPILL];
    trav_begin(st, ro, rd);
    if (WIDE && LOOP == LOOP_GATED) {
        trav4_gated<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack, live, leaf_gate);
    } else if (WIDE && LOOP == LOOP_POSTPONED) {
        trav4_postponed<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack, live);
    } else if (live) {
        if (WIDE && LOOP == LOOP_FUSED) {
            while (!trav_step4_fused<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack)) {
            }
        } else if (WIDE) {
            while (!trav_step4<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack)) {
            }
        } else {
            while (!trav_step<CLOSEST, BLOCK, LDS, STATS>(st, spill, nodes, tris, wr, s_stack)) {
            }
        }
    }
    return TraceOut{st.best_slot, st.best_t, st.nodes_visited, st.tris_tested};
}

// Tail hand-over.  A wave holds its SIMD slot until its slowest ray is done,
// and on sorted path-ray wavefronts the last ~15 of 64 rays take about a
// quarter of the wave's steps.  The compacting loop below therefore stops a
// wave once fewer than `compact` of its lanes are still walking, writes each
// walking ray's whole traversal state to the tail queue, and lets a tail
// pass (trace_tail_kernel) finish those rays in full waves.  The state is
// the ray slot, the current node, the stack (at most LDS_DEPTH entries, all
// in LDS; a wave with a deeper stack keeps walking) and the closest hit so
// far, so the tail pass continues exactly the same step sequence and the
// results do not change.  Record: 2 + LDS_DEPTH / 4 int4 = {slot, node, sp,
// best slot}, {best original index, best t, 0, 0}, then the LDS_DEPTH stack
// entries.
static_assert(LDS_DEPTH % 4 == 0, "tail record holds the LDS stack as int4");
constexpr int TAIL_INT4 = 2 + LDS_DEPTH / 4;

struct TailQueue {
    int4* q = nullptr;          // cap records
    uint32_t* ctr = nullptr;    // [0]: records reserved (may exceed cap)
    int64_t cap = 0;
    int compact = 0;            // hand over below this many walking lanes
};

// Closest-hit or any-hit walk of the 4-wide hierarchy with the single-load
// loop, handing the wave's walking rays over to `tq` once fewer than
// tq.compact lanes walk (one attempt per wave).  Returns true when this
// lane's ray was handed over (its record is then written by the tail pass).
template <bool CLOSEST, int BLOCK, int STACK>
__device__ __forceinline__ bool traverse4_handover(TravState<true, STACK>& st,
                                                   int (&spill)[TravState<true, STACK>::SPILL],
                                                   const float4* __restrict__ nodes4,
                                                   const float4* __restrict__ tris,
                                                   const WtRay& wr, bool live, int* s_stack,
                                                   int64_t slot, const TailQueue& tq) {
    bool active = live;
    bool tried = false;
    while (true) {
        const uint64_t act = ballot64(active);
        if (act == 0) return false;
        const int walking = __popcll(act);
        if (!tried && walking < tq.compact) {
            tried = true;
            if (__all(!active || st.sp <= LDS_DEPTH)) {
                const int leader = __ffsll((unsigned long long)act) - 1;
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(tq.ctr, (uint32_t)walking);
                base = __shfl(base, leader, 64);
                const int64_t rec = (int64_t)base + popc_below(act);
                if ((int64_t)base + walking <= tq.cap) {
                    if (!active) return false;
                    const lds_int* col = (const lds_int*)s_stack + threadIdx.x;
                    int4* o = tq.q + rec * TAIL_INT4;
                    o[0] = make_int4((int)slot, st.node, st.sp, st.best_slot);
                    o[1] = make_int4(st.best_orig, __float_as_int(st.best_t), 0, 0);
#pragma unroll
                    for (int k = 0; k < LDS_DEPTH / 4; ++k)
                        o[2 + k] = make_int4(col[(4 * k) * BLOCK], col[(4 * k + 1) * BLOCK],
                                             col[(4 * k + 2) * BLOCK], col[(4 * k + 3) * BLOCK]);
                    return true;
                }
                // The queue is full: void the reserved records below the
                // capacity and keep walking.
                if (active && rec < tq.cap) tq.q[rec * TAIL_INT4] = make_int4(-1, 0, 0, 0);
            }
        }
        if (active)
            active = !trav_step4_fused<CLOSEST, BLOCK, true, false>(st, spill, nodes4, tris, wr, s_stack);
    }
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
