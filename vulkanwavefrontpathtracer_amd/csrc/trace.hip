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
#include <mutex>

namespace wfpt {

// REORDER: the reorder gather fused in front of the trace (Tracer::traceSort
// with reorder, Tracer.cpp:209-225): ray_index is the sorted permutation,
// out ray i = rays[ray_index[i]] is written with its LUT entries exactly as
// reorder_kernel writes them, and traced from registers; its hit record
// goes to the reordered slot i.
// HANDOVER: the 4-wide single-load walk that hands the last rays of a wave
// to the tail queue `tq` (traverse4_handover); trace_tail_kernel finishes
// them.
template <bool CLOSEST, int BLOCK, bool LDS, bool STATS = false, bool WIDE = false,
          int LOOP = TRACE_LOOP_DEFAULT, bool REORDER = false, bool HANDOVER = false>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const float4* __restrict__ nodes,
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
                                                      int leaf_gate = 1,
                                                      TailQueue tq = TailQueue(), int xcd = 0) {
    static_assert(!HANDOVER || (LDS && WIDE && !STATS), "hand-over needs the 4-wide LDS walk");
    __shared__ int s_stack[LDS ? LDS_DEPTH * BLOCK : 1];
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
    if constexpr (HANDOVER) {
        TravState<true, TRACE_STACK4> st;
        int spill[TravState<true, TRACE_STACK4>::SPILL];
        trav_begin(st, ro, rd);
        if (traverse4_handover<CLOSEST, BLOCK>(st, spill, nodes, tris, wr, live, s_stack, r, tq))
            return;
        *reinterpret_cast<int4*>(hits + r) =
            hit_record(tris, wr, TraceOut{st.best_slot, st.best_t, 0u, 0u});
        return;
    }
    const TraceOut tr =
        traverse<CLOSEST, BLOCK, LDS, STATS, WIDE, LOOP>(nodes, tris, ro, rd, wr, live, s_stack,
                                                         leaf_gate);
    *reinterpret_cast<int4*>(hits + r) = hit_record(tris, wr, tr);
    if (STATS) {
        stat_nodes[r] = tr.nodes_visited;
        stat_tris[r] = tr.tris_tested;
    }
}

// Tail pass of the hand-over: a grid of resident waves finishes the rays in
// the tail queue.  Records are dealt statically (wave w of W takes the
// 64-record chunks w, w + W, ...) and a wave takes the next records of its
// sequence whenever `refill` or more of its lanes are idle.  A record
// restores the ray's traversal state (node, LDS stack, closest hit so far)
// and the walk continues with the same steps as in trace_kernel; the hit
// record goes to the slot the record names (rays[slot], hits[slot]).
// Voided records (slot -1, from a reservation that overran the queue) are
// skipped.
template <bool CLOSEST>
__global__ __launch_bounds__(64) void trace_tail_kernel(const float4* __restrict__ nodes4,
                                                        const float4* __restrict__ tris,
                                                        const wf_ray* __restrict__ rays,
                                                        wf_hit* __restrict__ hits,
                                                        const int4* __restrict__ q,
                                                        const uint32_t* __restrict__ ctr,
                                                        int64_t cap, int refill) {
    __shared__ int s_stack[LDS_DEPTH * 64];
    constexpr int STACK = TRACE_STACK4;
    const int64_t count = std::min<int64_t>((int64_t)ctr[0], cap);
    const int64_t num_waves = gridDim.x, wave = blockIdx.x;
    auto rec_of = [&](int64_t j) { return (((j >> 6) * num_waves + wave) << 6) + (j & 63); };
    if (rec_of(0) >= count) return;
    TravState<true, STACK> st;
    int spill[TravState<true, STACK>::SPILL];
    WtRay wr;
    int64_t r = 0, next = 0;
    bool active = false, exhausted = false;
    lds_int* col = (lds_int*)s_stack + threadIdx.x;
    while (true) {
        const uint64_t idle = ballot64(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && (nidle >= refill || nidle == 64)) {
            if (!active) {
                const int64_t k = rec_of(next + popc_below(idle));
                if (k < count) {
                    const int4* rec = q + k * TAIL_INT4;
                    const int4 h0 = rec[0];
                    if (h0.x >= 0) {
                        r = h0.x;
                        float4 ro, rd;
                        load_ray(rays, r, ro, rd);
                        (void)wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
                        trav_begin(st, ro, rd);
                        const int4 h1 = rec[1];
                        st.node = h0.y;
                        st.sp = h0.z;
                        st.best_slot = h0.w;
                        st.best_orig = h1.x;
                        st.best_t = __int_as_float(h1.y);
#pragma unroll
                        for (int e = 0; e < LDS_DEPTH / 4; ++e) {
                            const int4 v = rec[2 + e];
                            col[(4 * e) * 64] = v.x;
                            col[(4 * e + 1) * 64] = v.y;
                            col[(4 * e + 2) * 64] = v.z;
                            col[(4 * e + 3) * 64] = v.w;
                        }
                        active = true;
                    }
                }
            }
            next += nidle;
            exhausted = rec_of(next) >= count;
        }
        if (ballot64(active) == 0) {
            if (exhausted) break;
            continue;
        }
        if (active && trav_step4_fused<CLOSEST, 64, true, false>(st, spill, nodes4, tris, wr, s_stack)) {
            *reinterpret_cast<int4*>(hits + r) =
                hit_record(tris, wr, TraceOut{st.best_slot, st.best_t, 0u, 0u});
            active = false;
        }
    }
}

// Persistent variant: a grid sized to the resident waves; each wave keeps
// its lanes busy by handing new rays to idle lanes whenever `refill` or more
// are idle, so a wave is no longer held by its slowest of 64 rays (Aila &
// Laine 2009).  The rays are dealt statically: wave w of W owns the 64-ray
// chunks w, w + W, w + 2W, ... and takes them in order, so no counter (and
// no device-scope atomic) is involved, a chunk keeps the coherence of a
// sorted wavefront, and every wave samples the whole wavefront, which
// balances the work.  Each ray's traversal is the same step sequence as in
// trace_kernel, and the hit record fields are kept when a hit is accepted,
// so a finished lane stores its record without reloading the triangle.
template <bool CLOSEST, int BLOCK, bool LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) void trace_persistent_kernel(
    const float4* __restrict__ nodes, const float4* __restrict__ tris,
    const wf_ray* __restrict__ rays, const uint32_t* __restrict__ ray_index,
    wf_hit* __restrict__ hits, int64_t n, int refill) {
    __shared__ int s_stack[LDS ? LDS_DEPTH * BLOCK : 1];
    const unsigned lane = lane_id();
    constexpr int STACK = WIDE ? TRACE_STACK4 : TRACE_STACK;
    const int64_t num_waves = (int64_t)gridDim.x * (BLOCK / 64);
    const int64_t wave = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    // j-th ray of this wave's sequence (monotone in j)
    auto ray_of = [&](int64_t j) { return (((j >> 6) * num_waves + wave) << 6) + (j & 63); };
    TravState<LDS, STACK> st;
    int spill[TravState<LDS, STACK>::SPILL];
    WtRay wr;
    int64_t r = 0, next = 0;
    bool active = false, exhausted = ray_of(0) >= n;
    while (true) {
        const uint64_t idle = ballot64(!active);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= refill) {
            if (!active) {
                const int64_t k = ray_of(next + popc_below(idle));
                if (k < n) {
                    r = ray_index ? (int64_t)ray_index[k] : k;
                    float4 ro, rd;
                    load_ray(rays, r, ro, rd);
                    const bool live = (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
                    trav_begin(st, ro, rd);
                    if (live) active = true;
                    else *reinterpret_cast<int4*>(hits + r) = make_int4(-1, -1, __float_as_int(-1.0f), 0);
                }
            }
            next += nidle;
            exhausted = ray_of(next) >= n;
        }
        if (ballot64(active) == 0) {
            if (exhausted) break;
            continue;
        }
        bool fin = false;
        if (active) {
            if (WIDE) fin = trav_step4<CLOSEST, BLOCK, LDS, false, STACK, true>(st, spill, nodes, tris, wr, s_stack);
            else fin = trav_step<CLOSEST, BLOCK, LDS, false, STACK, true>(st, spill, nodes, tris, wr, s_stack);
        }
        if (fin) {
            *reinterpret_cast<int4*>(hits + r) =
                st.best_slot < 0 ? make_int4(-1, -1, __float_as_int(-1.0f), 0)
                                 : make_int4(st.best_prim, st.best_geom, __float_as_int(st.best_t),
                                             (int)st.best_bary);
            active = false;
        }
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
                const float4* nd = nodes4 + 4 * (int64_t)node;
                float tn[4];
                node4_hits(st, nd[0], nd[1], nd[2], nd[3], tn, ch);
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

constexpr int TRACE_MAX_DEVICES = 64;

template <bool CLOSEST, int BLOCK, bool LDS, bool WIDE>
static int launch_persistent(const float4* nodes, const float4* tris, const wf_ray* rays,
                             const uint32_t* ray_index, wf_hit* hits, int64_t n, int refill,
                             hipStream_t s) {
    static int resident[TRACE_MAX_DEVICES] = {};
    int dev = 0;
    WF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= TRACE_MAX_DEVICES) return fail(WF_ERR_HIP, "trace: device index out of range");
    if (resident[dev] == 0) {
        int per_cu = 0, cus = 0;
        WF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, trace_persistent_kernel<CLOSEST, BLOCK, LDS, WIDE>, BLOCK, 0));
        WF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        resident[dev] = std::max(1, per_cu) * std::max(1, cus);
    }
    const int64_t blocks = std::min<int64_t>(resident[dev], (n + BLOCK - 1) / BLOCK);
    trace_persistent_kernel<CLOSEST, BLOCK, LDS, WIDE><<<(unsigned)blocks, BLOCK, 0, s>>>(
        nodes, tris, rays, ray_index, hits, n, refill);
    return WF_OK;
}

template <bool CLOSEST, int BLOCK, bool LDS, bool WIDE>
static void launch_trace_loop(int loop, unsigned g, const float4* nodes, const float4* tris,
                              const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                              int64_t n, hipStream_t s, const TailQueue& tq) {
    if constexpr (WIDE && LDS) {
        if (loop == LOOP_FUSED && tq.q) {
            trace_kernel<CLOSEST, BLOCK, true, false, true, LOOP_FUSED, false, true><<<g, BLOCK, 0, s>>>(
                nodes, tris, rays, ray_index, hits, n, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, 1, tq, tuning().trace_xcd);
            return;
        }
    }
    if (WIDE && loop == LOOP_GATED)
        trace_kernel<CLOSEST, BLOCK, LDS, false, WIDE, LOOP_GATED><<<g, BLOCK, 0, s>>>(
            nodes, tris, rays, ray_index, hits, n, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, tuning().trace_leaf_gate);
    else if (WIDE && loop == LOOP_FUSED)
        trace_kernel<CLOSEST, BLOCK, LDS, false, WIDE, LOOP_FUSED><<<g, BLOCK, 0, s>>>(
            nodes, tris, rays, ray_index, hits, n, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, 1, TailQueue(), tuning().trace_xcd);
    else if (WIDE && loop == LOOP_POSTPONED)
        trace_kernel<CLOSEST, BLOCK, LDS, false, WIDE, LOOP_POSTPONED><<<g, BLOCK, 0, s>>>(nodes, tris, rays, ray_index, hits, n);
    else
        trace_kernel<CLOSEST, BLOCK, LDS, false, WIDE, LOOP_BRANCH><<<g, BLOCK, 0, s>>>(nodes, tris, rays, ray_index, hits, n);
}

template <bool CLOSEST, int BLOCK, bool WIDE>
static int launch_trace(bool lds, bool persistent, int loop, int refill, const float4* nodes,
                        const float4* tris, const wf_ray* rays, const uint32_t* ray_index,
                        wf_hit* hits, int64_t n, hipStream_t s, const TailQueue& tq) {
    if (persistent)
        return lds ? launch_persistent<CLOSEST, BLOCK, true, WIDE>(nodes, tris, rays, ray_index, hits, n, refill, s)
                   : launch_persistent<CLOSEST, BLOCK, false, WIDE>(nodes, tris, rays, ray_index, hits, n, refill, s);
    const unsigned g = grid_for(n, BLOCK);
    if (lds) launch_trace_loop<CLOSEST, BLOCK, true, WIDE>(loop, g, nodes, tris, rays, ray_index, hits, n, s, tq);
    else launch_trace_loop<CLOSEST, BLOCK, false, WIDE>(loop, g, nodes, tris, rays, ray_index, hits, n, s, tq);
    return WF_OK;
}

template <bool CLOSEST, bool WIDE>
static int launch_trace_block(const Tuning& tu, const float4* nodes, const float4* tris,
                              const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                              int64_t n, hipStream_t s, const TailQueue& tq) {
    const bool lds = tu.trace_lds_stack != 0, pers = tu.trace_persistent != 0;
    const int loop = tu.trace_loop;
    switch (tu.trace_block) {
        case 64: return launch_trace<CLOSEST, 64, WIDE>(lds, pers, loop, tu.trace_refill, nodes, tris, rays, ray_index, hits, n, s, tq);
        case 128: return launch_trace<CLOSEST, 128, WIDE>(lds, pers, loop, tu.trace_refill, nodes, tris, rays, ray_index, hits, n, s, tq);
        default: return launch_trace<CLOSEST, 256, WIDE>(lds, pers, loop, tu.trace_refill, nodes, tris, rays, ray_index, hits, n, s, tq);
    }
}

// Tail queue of a trace of n rays on `scene`: set (q non-null) when the
// hand-over is on and applies (4-wide single-load LDS walk, n at least
// trace.compact_min), with its counter cleared on stream s.  The queue is
// allocated on first use (trace.tail_cap records of 16 * TAIL_INT4 bytes, by default
// max(2^20, n/2), kept with the scene); a first use under stream capture
// walks without hand-over instead of allocating.
static std::mutex g_tailq_mutex;

static int tail_queue(wf_scene sc, int64_t n, hipStream_t s, TailQueue& tq) {
    const Tuning& tu = tuning();
    tq = TailQueue();
    if (tu.trace_compact <= 0 || n < tu.trace_compact_min || n > INT32_MAX || tu.trace_width != 4 ||
        !tu.trace_lds_stack || tu.trace_persistent || tu.trace_loop != LOOP_FUSED)
        return WF_OK;
    {
        std::lock_guard<std::mutex> lock(g_tailq_mutex);
        if (!sc->tailq) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            WF_HIP(hipStreamIsCapturing(s, &cs));
            if (cs != hipStreamCaptureStatusNone) return WF_OK;
            const int64_t cap = tu.trace_tail_cap > 0 ? (int64_t)tu.trace_tail_cap
                                                      : std::max<int64_t>(int64_t(1) << 20, n / 2);
            int4* q = nullptr;
            uint32_t* c = nullptr;
            if (hipMalloc((void**)&q, sizeof(int4) * TAIL_INT4 * cap) != hipSuccess)
                return fail(WF_ERR_OUT_OF_MEMORY, "trace: tail queue allocation failed");
            if (hipMalloc((void**)&c, 4 * sizeof(uint32_t)) != hipSuccess) {
                (void)hipFree(q);
                return fail(WF_ERR_OUT_OF_MEMORY, "trace: tail queue allocation failed");
            }
            sc->tailq = q;
            sc->tailq_ctr = c;
            sc->tailq_cap = cap;
        }
    }
    WF_TRY(fill_u32(sc->tailq_ctr, 1, 0u, s));
    tq.q = sc->tailq;
    tq.ctr = sc->tailq_ctr;
    tq.cap = sc->tailq_cap;
    tq.compact = tu.trace_compact;
    return WF_OK;
}

// The tail pass over the queue filled by the preceding hand-over trace.
template <bool CLOSEST>
static int launch_tail(const float4* nodes4, const float4* tris, const wf_ray* rays, wf_hit* hits,
                       const TailQueue& tq, hipStream_t s) {
    static int resident[TRACE_MAX_DEVICES] = {};
    int dev = 0;
    WF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= TRACE_MAX_DEVICES) return fail(WF_ERR_HIP, "trace: device index out of range");
    if (resident[dev] == 0) {
        int per_cu = 0, cus = 0;
        WF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_tail_kernel<CLOSEST>, 64, 0));
        WF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        resident[dev] = std::max(1, per_cu) * std::max(1, cus);
    }
    const int64_t waves = std::min<int64_t>(resident[dev], (tq.cap + 63) / 64);
    trace_tail_kernel<CLOSEST><<<(unsigned)waves, 64, 0, s>>>(nodes4, tris, rays, hits, tq.q, tq.ctr,
                                                              tq.cap, tuning().trace_tail_refill);
    return launch_check("trace_tail_kernel");
}

template <bool CLOSEST, int BLOCK>
static void launch_reorder_trace_block(int loop, const float4* nodes4, const float4* tris,
                                       const uint32_t* idx, const wf_ray* in_rays,
                                       const int32_t* in_s2i, wf_ray* out_rays, int32_t* out_s2i,
                                       int32_t* out_i2s, wf_hit* hits, int64_t n, hipStream_t s,
                                       const TailQueue& tq) {
    const unsigned g = grid_for(n, BLOCK);
    if (loop == LOOP_FUSED && tq.q)
        trace_kernel<CLOSEST, BLOCK, true, false, true, LOOP_FUSED, true, true><<<g, BLOCK, 0, s>>>(
            nodes4, tris, in_rays, idx, hits, n, nullptr, nullptr, in_s2i, out_rays, out_s2i, out_i2s,
            1, tq, tuning().trace_xcd);
    else if (loop == LOOP_GATED)
        trace_kernel<CLOSEST, BLOCK, true, false, true, LOOP_GATED, true><<<g, BLOCK, 0, s>>>(
            nodes4, tris, in_rays, idx, hits, n, nullptr, nullptr, in_s2i, out_rays, out_s2i, out_i2s,
            tuning().trace_leaf_gate);
    else if (loop == LOOP_FUSED)
        trace_kernel<CLOSEST, BLOCK, true, false, true, LOOP_FUSED, true><<<g, BLOCK, 0, s>>>(
            nodes4, tris, in_rays, idx, hits, n, nullptr, nullptr, in_s2i, out_rays, out_s2i, out_i2s,
            1, TailQueue(), tuning().trace_xcd);
    else
        trace_kernel<CLOSEST, BLOCK, true, false, true, LOOP_BRANCH, true><<<g, BLOCK, 0, s>>>(
            nodes4, tris, in_rays, idx, hits, n, nullptr, nullptr, in_s2i, out_rays, out_s2i, out_i2s);
}

// The fused reorder + trace kernel serves the one-lane, LDS-stack, 4-wide
// traversal (branch or single-load loop); other configurations run the
// reorder kernel and then the trace.  *fused is false when not fused.
static int launch_reorder_trace(wf_scene scene, const uint32_t* idx, const wf_ray* in_rays,
                                const int32_t* in_s2i, wf_ray* out_rays, int32_t* out_s2i,
                                int32_t* out_i2s, wf_hit* hits, int64_t n, int closest,
                                hipStream_t s, bool* fused) {
    const Tuning& tu = tuning();
    const bool grouped = n <= tu.trace_group64 || n <= tu.trace_group16 || n <= tu.trace_group4;
    *fused = false;
    if (!tu.trace_fuse_reorder || scene->num_tris == 0 || tu.trace_width != 4 ||
        tu.trace_persistent || !tu.trace_lds_stack ||
        grouped || tu.trace_loop == LOOP_POSTPONED)
        return WF_OK;
    *fused = true;
    TailQueue tq;
    WF_TRY(tail_queue(scene, n, s, tq));
    const float4 *nodes4 = scene->nodes4, *tris = scene->tris;
    switch (tu.trace_block) {
        case 64:
            if (closest) launch_reorder_trace_block<true, 64>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            else launch_reorder_trace_block<false, 64>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            break;
        case 128:
            if (closest) launch_reorder_trace_block<true, 128>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            else launch_reorder_trace_block<false, 128>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            break;
        default:
            if (closest) launch_reorder_trace_block<true, 256>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            else launch_reorder_trace_block<false, 256>(tu.trace_loop, nodes4, tris, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n, s, tq);
            break;
    }
    WF_TRY(launch_check("trace_kernel<reorder>"));
    if (tq.q) {
        // the handed-over rays were written to out_rays[i] before their walk
        if (closest) WF_TRY(launch_tail<true>(nodes4, tris, out_rays, hits, tq, s));
        else WF_TRY(launch_tail<false>(nodes4, tris, out_rays, hits, tq, s));
    }
    return WF_OK;
}

int trace_launch(wf_scene scene, const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                 int64_t n, int closest, hipStream_t s) {
    if (n == 0) return WF_OK;
    if (scene->num_tris == 0) {
        miss_kernel<<<grid_for(n, 256), 256, 0, s>>>(ray_index, hits, n);
        return launch_check("miss_kernel");
    }
    const Tuning& tu = tuning();
    if (tu.trace_width == 4 && !tu.trace_persistent &&
        (closest ? launch_group<true>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s)
                 : launch_group<false>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s)))
        return launch_check("trace_group_kernel");
    TailQueue tq;
    WF_TRY(tail_queue(scene, n, s, tq));
    if (tu.trace_width == 4) {
        if (closest)
            WF_TRY((launch_trace_block<true, true>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s, tq)));
        else
            WF_TRY((launch_trace_block<false, true>(tu, scene->nodes4, scene->tris, rays, ray_index, hits, n, s, tq)));
    } else {
        if (closest)
            WF_TRY((launch_trace_block<true, false>(tu, scene->nodes, scene->tris, rays, ray_index, hits, n, s, tq)));
        else
            WF_TRY((launch_trace_block<false, false>(tu, scene->nodes, scene->tris, rays, ray_index, hits, n, s, tq)));
    }
    WF_TRY(launch_check("trace_kernel"));
    if (tq.q) {
        // handed-over records name the slot r of rays[r] / hits[r]
        if (closest) WF_TRY(launch_tail<true>(scene->nodes4, scene->tris, rays, hits, tq, s));
        else WF_TRY(launch_tail<false>(scene->nodes4, scene->tris, rays, hits, tq, s));
    }
    return WF_OK;
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
    const bool wide = tuning().trace_width == 4;
    const float4* nodes = wide ? scene->nodes4 : scene->nodes;
    const int loop = tuning().trace_loop;
    if (closest && wide && loop == LOOP_BRANCH)
        trace_kernel<true, 64, true, true, true, LOOP_BRANCH><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else if (closest && wide && loop == LOOP_POSTPONED)
        trace_kernel<true, 64, true, true, true, LOOP_POSTPONED><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else if (closest && wide)
        trace_kernel<true, 64, true, true, true><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else if (closest)
        trace_kernel<true, 64, true, true, false><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else if (wide)
        trace_kernel<false, 64, true, true, true><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else
        trace_kernel<false, 64, true, true, false><<<grid_for(n, 64), 64, 0, s>>>(
            nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
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
    bool fused = false;
    WF_TRY(launch_reorder_trace(scene, idx, in_rays, in_s2i, out_rays, out_s2i, out_i2s, hits, n,
                                closest, s, &fused));
    if (!fused) {
        WF_TRY(wf_reorder(idx, in_rays, out_rays, in_s2i, out_s2i, out_i2s, n, stream, nullptr));
        WF_TRY(trace_launch(scene, out_rays, nullptr, hits, n, closest, s));
    }
    return t.finish();
}
