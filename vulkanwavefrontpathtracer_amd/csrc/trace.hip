// Ray traversal: replaces trace.comp:27-88 (rayQuery against the driver's
// acceleration structure) with a stack-based traversal of the LBVH, one ray
// per lane of a wave64.
//
// Semantics kept from the reference:
//   - opaque geometry, no face culling (TLAS instance flag
//     TRIANGLE_FACING_CULL_DISABLE, WaveFrontPathTracer.cpp:367);
//   - interval [tmin, tmax] of the ray; tmax < tmin is a miss (dead shadow
//     rays carry tmax = -1, raygenShadow.comp:129);
//   - closest hit, or any hit with early termination (trace.comp:36-37);
//   - optional indirection r = ray_index[i], result written to hits[r]
//     (trace.comp:43-44, :84);
//   - hit record {prim in geometry, geometry, t, packUnorm2x16(bary)} and
//     miss record {-1, -1, -1, 0} (trace.comp:70-82).
// Closest-hit ties on t resolve to the smaller original triangle index, so
// the result equals a brute-force scan whatever the traversal order.
#include "traverse.h"

#include <algorithm>

namespace wfpt {

// REORDER: the reorder gather fused in front of the trace (Tracer::traceSort
// with reorder, Tracer.cpp:209-225): ray_index is the sorted permutation,
// out ray i = rays[ray_index[i]] is written with its LUT entries exactly as
// reorder_kernel writes them, and traced from registers; its hit record
// goes to the reordered slot i.  xcd > 0 hands runs of xcd consecutive
// blocks to each XCD (see xcd_block).
template <bool CLOSEST, int BLOCK, bool STATS = false, bool REORDER = false>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const float4* __restrict__ nodes4,
                                                      const float4* __restrict__ tris,
                                                      const wf_ray* __restrict__ rays,
                                                      const uint32_t* __restrict__ ray_index,
                                                      wf_hit* __restrict__ hits, int64_t n,
                                                      uint32_t* __restrict__ stat_nodes = nullptr,
                                                      uint32_t* __restrict__ stat_tris = nullptr,
                                                      const int32_t* __restrict__ in_s2i = nullptr,
                                                      wf_ray* __restrict__ out_rays = nullptr,
                                                      int32_t* __restrict__ out_s2i = nullptr,
                                                      int32_t* __restrict__ out_i2s = nullptr,
                                                      int xcd = 0) {
    __shared__ int s_stack[LDS_DEPTH * BLOCK];
    const int64_t blk = xcd ? xcd_block(blockIdx.x, gridDim.x, xcd) : (int64_t)blockIdx.x;
    int64_t i = blk * BLOCK + threadIdx.x;
    if (i >= n) return;
    int64_t r;
    float4 ro, rd;
    if (REORDER) {
        const int64_t src = ray_index[i];
        load_ray(rays, src, ro, rd);
        store_ray(out_rays, i, ro, rd);
        const int32_t index = in_s2i[src];
        out_s2i[i] = index;
        out_i2s[index] = (int32_t)i;
        r = i;
    } else {
        r = ray_index ? (int64_t)ray_index[i] : i;
        load_ray(rays, r, ro, rd);
    }
    WtRay wr;
    wr.uperm = -1;
    const bool live = (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
    wt_uniform_perm(wr, live);
    const TraceOut tr = traverse<CLOSEST, BLOCK, STATS>(nodes4, tris, ro, rd, wr, live, s_stack);
    *reinterpret_cast<int4*>(hits + r) = hit_record(tris, wr, tr);
    if (STATS) {
        stat_nodes[r] = tr.nodes_visited;
        stat_tris[r] = tr.tris_tested;
    }
}

// Small wavefronts: a group of G lanes (4, 16 or 64) traces one ray of the
// 4-wide hierarchy.  The group expands the top levels breadth first (while
// the next level fits in G lanes) into at most G subtrees, then each lane
// walks one subtree with its own stack, and after every step the lanes take
// the group's closest hit distance for culling.  The last bounces of a frame
// have a few thousand rays, and a launch then lasts as long as the chain of
// dependent node fetches of its longest ray (about 150 steps for a ray that
// crosses a whole room); spreading the walk over G lanes shortens that
// chain.  The record is the group's minimum of (t, original triangle index)
// over its lanes' hits, the same as a one-lane traversal's closest hit; an
// any-hit group stops at the first hit of any lane.  A lane that adopts a
// strictly closer group distance drops its own hit (it cannot win), so a
// later tie at that distance is decided by the original index alone.
template <bool CLOSEST, int G>
__global__ __launch_bounds__(64) void trace_group_kernel(const float4* __restrict__ nodes4,
                                                         const float4* __restrict__ tris,
                                                         const wf_ray* __restrict__ rays,
                                                         const uint32_t* __restrict__ ray_index,
                                                         wf_hit* __restrict__ hits, int64_t n) {
    static_assert(G == 4 || G == 16 || G == 64, "group of 4, 16 or 64 lanes");
    constexpr int STACK = TRACE_STACK4;
    __shared__ int s_stack[LDS_DEPTH * 64];
    // frontier of each group: up to FRONT nodes at group offset gbase * FRONT / G
    constexpr int FRONT = 4 * G;
    __shared__ int s_front[2][64 * 4];
    const unsigned lane = lane_id();
    const int gl = (int)(lane & (G - 1));
    const int gbase = (int)(lane & ~(unsigned)(G - 1));
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const int64_t ray = ((int64_t)blockIdx.x * 64 + lane) / G;
    const bool valid = ray < n;
    int64_t r = 0;
    float4 ro = make_float4(0.f, 0.f, 0.f, 0.f), rd = make_float4(0.f, 0.f, 1.f, -1.f);
    if (valid) {
        r = ray_index ? (int64_t)ray_index[ray] : ray;
        load_ray(rays, r, ro, rd);
    }
    WtRay wr;
    wr.uperm = -1;
    const bool live = valid && (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
    wt_uniform_perm(wr, live);
    TravState<true, STACK> st;
    int spill[TravState<true, STACK>::SPILL];
    trav_begin(st, ro, rd);

    // Breadth-first expansion: before each level the frontier holds at most
    // G nodes (one per lane), so the next one (at most 4 children each) fits
    // in the group's 4G frontier slots; the last level leaves up to 4G
    // subtrees, which the lanes then take from the frontier one at a time.
    const int fbase = gbase * 4;
    int cur = 0, fsize = live ? 1 : 0;
    if (gl == 0) s_front[0][fbase] = 0;
    __syncthreads();
#pragma unroll
    for (int level = 0; (1 << (2 * level)) <= G; ++level) {
        int c = 0;
        int ch[4] = {0, 0, 0, 0};
        if (gl < fsize) {
            const int node = s_front[cur][fbase + gl];
            if (node >= 0) {
                const float4* nd = nodes4 + NODE4_F4 * (int64_t)node;
                float tn[4];
                node4_hits(st, nd[0], nd[1], nd[2], nd[3], nd[4], tn, ch);
                c = (tn[0] < INFINITY) + (tn[1] < INFINITY) + (tn[2] < INFINITY) + (tn[3] < INFINITY);
            } else {
                ch[0] = node;  // a leaf stays in the frontier
                c = 1;
            }
        }
        int incl = c;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int y = __shfl_up(incl, o, G);
            if (gl >= o) incl += y;
        }
        const int total = __shfl(incl, G - 1, G);
        for (int k = 0; k < c; ++k) s_front[cur ^ 1][fbase + incl - c + k] = ch[k];
        __syncthreads();
        cur ^= 1;
        fsize = total;
    }
    static_assert(FRONT <= 256, "frontier slots of a wave");

    // Each lane walks one frontier subtree with its own stack; a lane whose
    // walk ends takes the group's next subtree, so the group's time is its
    // total work over G lanes plus one subtree, not its largest lane's share.
    bool active = false;
    int head = 0;  // next frontier subtree of the group (same on every lane)
    while (true) {
        const uint64_t want = ballot64(!active) & gmask;
        const int k = (int)popc_below(want);
        if (!active && head + k < fsize) {
            st.node = s_front[cur][fbase + head + k];
            active = true;
        }
        head = min(fsize, head + __popcll(want));
        if (!__any(active)) break;
        if (active)
            active = !trav_step4_fused<CLOSEST, 64, true, false, STACK, true>(st, spill, nodes4, tris, wr,
                                                                          s_stack);
        float bt = st.best_t;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) bt = fminf(bt, __shfl_xor(bt, o, G));
        if (bt < st.best_t) {
            st.best_t = bt;
            st.best_slot = -1;
            st.best_orig = 0x7fffffff;
        }
        if (!CLOSEST && (ballot64(st.best_slot >= 0) & gmask)) {
            active = false;
            head = fsize;
        }
    }

    // group minimum of (t, original index) over the lanes that hold a hit
    const bool has = st.best_slot >= 0;
    const float t = has ? st.best_t : INFINITY;
    const int o = has ? st.best_orig : 0x7fffffff;
    float tm = t;
    int om = o;
#pragma unroll
    for (int sft = 1; sft < G; sft <<= 1) {
        const float t2 = __shfl_xor(tm, sft, G);
        const int o2 = __shfl_xor(om, sft, G);
        if (t2 < tm || (t2 == tm && o2 < om)) {
            tm = t2;
            om = o2;
        }
    }
    const uint64_t winners = ballot64(has && t == tm && o == om) & gmask;
    if (valid) {
        if (winners == 0) {
            if (gl == 0) *reinterpret_cast<int4*>(hits + r) = make_int4(-1, -1, __float_as_int(-1.0f), 0);
        } else if ((int)lane == __ffsll((unsigned long long)winners) - 1) {
            *reinterpret_cast<int4*>(hits + r) =
                make_int4(st.best_prim, st.best_geom, __float_as_int(st.best_t), (int)st.best_bary);
        }
    }
}

template <bool CLOSEST>
static bool launch_group(const Tuning& tu, const float4* nodes4, const float4* tris,
                         const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits, int64_t n,
                         hipStream_t s) {
    if (n <= tu.trace_group64)
        trace_group_kernel<CLOSEST, 64><<<grid_for(n * 64, 64), 64, 0, s>>>(nodes4, tris, rays, ray_index, hits, n);
    else if (n <= tu.trace_group16)
        trace_group_kernel<CLOSEST, 16><<<grid_for(n * 16, 64), 64, 0, s>>>(nodes4, tris, rays, ray_index, hits, n);
    else if (n <= tu.trace_group4)
        trace_group_kernel<CLOSEST, 4><<<grid_for(n * 4, 64), 64, 0, s>>>(nodes4, tris, rays, ray_index, hits, n);
    else
        return false;
    return true;
}

__global__ __launch_bounds__(256) void miss_kernel(const uint32_t* __restrict__ ray_index,
                                                   wf_hit* __restrict__ hits, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t r = ray_index ? (int64_t)ray_index[i] : i;
    *reinterpret_cast<int4*>(hits + r) = make_int4(-1, -1, __float_as_int(-1.0f), 0);
}

template <bool CLOSEST, int BLOCK, bool REORDER>
static void launch_trace_kernel(const float4* nodes4, const float4* tris, const wf_ray* rays,
                                const uint32_t* ray_index, wf_hit* hits, int64_t n,
                                const int32_t* in_s2i, wf_ray* out_rays, int32_t* out_s2i,
                                int32_t* out_i2s, hipStream_t s) {
    trace_kernel<CLOSEST, BLOCK, false, REORDER><<<grid_for(n, BLOCK), BLOCK, 0, s>>>(
        nodes4, tris, rays, ray_index, hits, n, nullptr, nullptr, in_s2i, out_rays, out_s2i,
        out_i2s, tuning().trace_xcd);
}

// One-lane-per-ray trace (REORDER: with the reorder gather fused in front)
// at the configured block size.
template <bool REORDER>
static void launch_trace_block(int closest, const float4* nodes4, const float4* tris,
                               const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                               int64_t n, const int32_t* in_s2i, wf_ray* out_rays,
                               int32_t* out_s2i, int32_t* out_i2s, hipStream_t s) {
    switch (tuning().trace_block) {
        case 64:
            if (closest) launch_trace_kernel<true, 64, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            else launch_trace_kernel<false, 64, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            break;
        case 128:
            if (closest) launch_trace_kernel<true, 128, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            else launch_trace_kernel<false, 128, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            break;
        default:
            if (closest) launch_trace_kernel<true, 256, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            else launch_trace_kernel<false, 256, REORDER>(nodes4, tris, rays, ray_index, hits, n, in_s2i, out_rays, out_s2i, out_i2s, s);
            break;
    }
}

static bool grouped_size(int64_t n) {
    const Tuning& tu = tuning();
    return n <= tu.trace_group64 || n <= tu.trace_group16 || n <= tu.trace_group4;
}

int trace_launch(wf_scene scene, const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                 int64_t n, int closest, hipStream_t s) {
    if (n == 0) return WF_OK;
    if (scene->num_tris == 0) {
        miss_kernel<<<grid_for(n, 256), 256, 0, s>>>(ray_index, hits, n);
        return launch_check("miss_kernel");
    }
    const Tuning& tu = tuning();
    if (closest ? launch_group<true>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s)
                : launch_group<false>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s))
        return launch_check("trace_group_kernel");
    launch_trace_block<false>(closest, scene->nodes4, scene->tris, rays, ray_index, hits, n, nullptr,
                              nullptr, nullptr, nullptr, s);
    return launch_check("trace_kernel");
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_trace_stats(wf_scene scene, const wf_ray* rays, wf_hit* hits, int64_t n,
                              int closest, uint32_t* node_visits, uint32_t* tri_tests,
                              wf_stream_t stream) {
    WF_CHECK_ARG(scene && n >= 0, "wf_trace_stats: bad scene or n");
    if (n == 0) return WF_OK;
    WF_CHECK_ARG(rays && hits && node_visits && tri_tests, "wf_trace_stats: null buffer");
    WF_CHECK_ARG(scene->num_tris > 0, "wf_trace_stats: empty scene");
    hipStream_t s = as_stream(stream);
    if (closest)
        trace_kernel<true, 64, true><<<grid_for(n, 64), 64, 0, s>>>(
            scene->nodes4, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else
        trace_kernel<false, 64, true><<<grid_for(n, 64), 64, 0, s>>>(
            scene->nodes4, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    return launch_check("trace_kernel<stats>");
}

extern "C" int wf_trace(wf_scene scene, const wf_ray* rays, const uint32_t* ray_index,
                        wf_hit* hits, int64_t n, int closest, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(scene, "wf_trace: null scene");
    WF_CHECK_ARG(n >= 0, "wf_trace: n < 0");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(rays && hits, "wf_trace: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    WF_TRY(trace_launch(scene, rays, ray_index, hits, n, closest, s));
    return t.finish();
}

extern "C" int wf_reorder_trace(wf_scene scene, const uint32_t* idx, const wf_ray* in_rays,
                                wf_ray* out_rays, const int32_t* in_s2i, int32_t* out_s2i,
                                int32_t* out_i2s, wf_hit* hits, int64_t n, int closest,
                                wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(scene, "wf_reorder_trace: null scene");
    WF_CHECK_ARG(n >= 0, "wf_reorder_trace: n < 0");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(idx && in_rays && out_rays && in_s2i && out_s2i && out_i2s && hits,
                 "wf_reorder_trace: null buffer");
    WF_CHECK_ARG((const void*)in_rays != (const void*)out_rays, "wf_reorder_trace: in_rays == out_rays");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    // the fused kernel serves the one-lane traversal; grouped sizes (and
    // trace.fuse_reorder 0) run the reorder and then the trace
    if (tuning().trace_fuse_reorder && scene->num_tris > 0 && !grouped_size(n)) {
        launch_trace_block<true>(closest, scene->nodes4, scene->tris, in_rays, idx, hits, n, in_s2i,
                                 out_rays, out_s2i, out_i2s, s);
        WF_TRY(launch_check("trace_kernel<reorder>"));
    } else {
        WF_TRY(wf_reorder(idx, in_rays, out_rays, in_s2i, out_s2i, out_i2s, n, stream, nullptr));
        WF_TRY(trace_launch(scene, out_rays, nullptr, hits, n, closest, s));
    }
    return t.finish();
}
