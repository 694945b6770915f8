// Stack-based BVH traversal of one ray by one lane of a wave64, shared by
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

// Plane value of child k from a pair of fp16 words (children 0/1 in w01,
// 2/3 in w23, the first of each pair in the low half).  Converted inside the
// fused multiply-add that uses it (v_fma_mix_f32 takes fp16 halves).
__device__ __forceinline__ float plane_of(uint32_t w01, uint32_t w23, int k) {
    const uint32_t w = k < 2 ? w01 : w23;
    const unsigned short h = (k & 1) ? (unsigned short)(w >> 16) : (unsigned short)(w & 0xffffu);
    return (float)__builtin_bit_cast(_Float16, h);
}

// Interior node of the 4-wide quantized hierarchy (header a, x/y/z plane
// words px/py/pz, references r; layout in scene.h): the four child boxes
// decoded from their planes (t = q * 2^e / d + (origin - o) / d, one fused
// multiply-add per plane) and tested against [tmin, best_t].  tn[k] is the
// entry distance of the k-th nearest hit child (INFINITY for a miss),
// ref[k] its reference.  Without FULL_ORDER (any-hit queries, which end at
// the first hit found whatever the order) only slot 0 is sorted: it holds the
// nearest hit child, or a miss when no child is hit.
template <bool LDS, int STACK, bool FULL_ORDER = true>
__device__ __forceinline__ void node4_hits(const TravState<LDS, STACK>& st, float4 a, float4 px,
                                           float4 py, float4 pz, float4 r, float (&tn)[4],
                                           int (&ref)[4]) {
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
    const bool fx = st.ix < 0.0f, fy = st.iy < 0.0f, fz = st.iz < 0.0f;
    const uint32_t nx01 = __float_as_uint(fx ? px.z : px.x), nx23 = __float_as_uint(fx ? px.w : px.y);
    const uint32_t fx01 = __float_as_uint(fx ? px.x : px.z), fx23 = __float_as_uint(fx ? px.y : px.w);
    const uint32_t ny01 = __float_as_uint(fy ? py.z : py.x), ny23 = __float_as_uint(fy ? py.w : py.y);
    const uint32_t fy01 = __float_as_uint(fy ? py.x : py.z), fy23 = __float_as_uint(fy ? py.y : py.w);
    const uint32_t nz01 = __float_as_uint(fz ? pz.z : pz.x), nz23 = __float_as_uint(fz ? pz.w : pz.y);
    const uint32_t fz01 = __float_as_uint(fz ? pz.x : pz.z), fz23 = __float_as_uint(fz ? pz.y : pz.w);
    ref[0] = __float_as_int(r.x);
    ref[1] = __float_as_int(r.y);
    ref[2] = __float_as_int(r.z);
    ref[3] = __float_as_int(r.w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float nx = fmaf(plane_of(nx01, nx23, k), sax, bx);
        const float fxv = fmaf(plane_of(fx01, fx23, k), sax, bx);
        const float ny = fmaf(plane_of(ny01, ny23, k), say, by);
        const float fyv = fmaf(plane_of(fy01, fy23, k), say, by);
        const float nz = fmaf(plane_of(nz01, nz23, k), saz, bz);
        const float fzv = fmaf(plane_of(fz01, fz23, k), saz, bz);
        const float nn = fmaxf(fmaxf(nx, ny), fmaxf(nz, st.tmin));
        const float ff = fminf(fminf(fxv, fyv), fminf(fzv, st.best_t));
        tn[k] = nn <= ff ? nn : INFINITY;
    }
    cswap(tn[0], ref[0], tn[1], ref[1]);
    cswap(tn[2], ref[2], tn[3], ref[3]);
    cswap(tn[0], ref[0], tn[2], ref[2]);
    if (FULL_ORDER) {
        cswap(tn[1], ref[1], tn[3], ref[3]);
        cswap(tn[1], ref[1], tn[2], ref[2]);
    }
}

// TOP: `top` holds the stack's top entry read before this node's pushes (the
// fused step reads it beside the node fetch); a node that pushes nothing and
// pops takes it instead of reading LDS again.
template <int BLOCK, bool LDS, bool STATS, int STACK, bool TOP = false, bool FULL_ORDER = true>
__device__ __forceinline__ bool trav_node4_data(TravState<LDS, STACK>& st,
                                                int (&spill)[TravState<LDS, STACK>::SPILL],
                                                float4 a, float4 px, float4 py, float4 pz, float4 r,
                                                int* s_stack, int top = 0) {
    if (STATS) ++st.nodes_visited;
    float tn[4];
    int ref[4];
    node4_hits<LDS, STACK, FULL_ORDER>(st, a, px, py, pz, r, tn, ref);
    const int h1 = tn[1] < INFINITY, h2 = tn[2] < INFINITY, h3 = tn[3] < INFINITY;
    if (LDS && __all(st.sp <= LDS_DEPTH - 3)) {
        // Three unconditional LDS writes, each at the top plus the hits
        // written before it, so the hit children land contiguously whatever
        // their order: a missed child's slot is overwritten by the next write
        // or left above the new top.
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

// One traversal step of the 4-wide hierarchy that issues a single load per
// lane whatever the lane visits: the node's 80 bytes, or the first triangle
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
    const float4* p = inner ? nodes4 + NODE4_F4 * (int64_t)st.node
                            : tris + 3 * (int64_t)leaf_first(st.node);
    const float4 a = p[0], b = p[1], c = p[2];
    // the node's last two quads; a leaf step never reads them, so they are
    // left unset there (no register moves on the leaf lanes)
    float4 d, e;
    if (inner) {
        d = p[3];
        e = p[4];
    }
    // The stack's top entry, read while the fetch is in flight: the next node
    // of a step that pushes nothing (a leaf, or a node with no child hit),
    // which then waits for neither LDS nor the fetch's dependents to pop.
    int top = 0;
    if (LDS && st.sp > 0) top = ((const lds_int*)s_stack)[min(st.sp - 1, LDS_DEPTH - 1) * BLOCK + threadIdx.x];
    if (inner)
        return !trav_node4_data<BLOCK, LDS, STATS, STACK, true, CLOSEST>(st, spill, a, b, c, d, e, s_stack,
                                                                          top);
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


// Closest-hit or any-hit walk of one ray through the 4-wide quantized
// hierarchy (nodes4) with the single-load step; s_stack is this block's LDS
// stack (LDS_DEPTH * BLOCK ints).
template <bool CLOSEST, int BLOCK, bool STATS>
__device__ __forceinline__ TraceOut traverse(const float4* __restrict__ nodes4,
                                             const float4* __restrict__ tris, float4 ro, float4 rd,
                                             const WtRay& wr, bool live, int* s_stack) {
    TravState<true, TRACE_STACK4> st;
    int spill[TravState<true, TRACE_STACK4>::SPILL];
    trav_begin(st, ro, rd);
    if (live) {
        while (!trav_step4_fused<CLOSEST, BLOCK, true, STATS>(st, spill, nodes4, tris, wr, s_stack)) {
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
