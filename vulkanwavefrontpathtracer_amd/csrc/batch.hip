// Kernels of the batched wavefront queue (batch.h).  Each restates the
// single-view stage of the same name with per-ray view and scene lookup and
// with the ray count read from the device, so a whole pass is a fixed launch
// sequence:
//   primary rays -> trace -> smooth -> shadow rays -> trace (any) -> shadow
//   then per bounce: shade/extend (compaction) -> [keys -> sort -> reorder]
//   -> trace -> smooth -> shadow rays -> trace (any) -> shadow.
// Per pixel the arithmetic is that of the single-view stages (same seeds,
// keyed by the view's own pixel index, same operation order), so every
// view's frame equals the single-view frame loop's bit for bit.
#include "batch.h"
#include "keys.h"
#include "traverse.h"

#include <cfloat>

namespace wfpt {

int sort_pairs_launch(void* keys, uint32_t* vals, int64_t n_cap, const uint32_t* d_n,
                      int key_bytes, int key_bits, int digit_bits, void* scratch,
                      int64_t scratch_bytes, hipStream_t s);

// View owning global pixel gp: offsets[v] <= gp < offsets[v + 1].
__device__ __forceinline__ int view_of(const BatchTables& t, int64_t gp) {
    int lo = 0, hi = t.num_views;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t.offsets[mid] <= gp) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t count_at(const uint32_t* q, int b) { return q[BQ_COUNT0 + b]; }

// ------------------------------------------------------------ frame / pass
__global__ __launch_bounds__(256) void batch_seeds_kernel(const BatchTables t,
                                                          uint32_t* __restrict__ seeds,
                                                          int64_t total, uint32_t frame) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int v = view_of(t, i);
    seeds[i] = tea((uint32_t)(i - t.offsets[v]), frame, 16);  // initSeeds.comp:31-42
}

__global__ void batch_set_pass_kernel(uint32_t* __restrict__ q, int frame, int pass,
                                      uint32_t count, int depth) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    q[BQ_PASS] = (uint32_t)pass;
    q[BQ_FRAME] = (uint32_t)frame;
    q[BQ_COUNT0] = count;
    for (int b = 1; b <= depth + 1; ++b) q[BQ_COUNT0 + b] = 0u;
}

__global__ void batch_totals_kernel(uint32_t* __restrict__ q, int depth, int S) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t path = 0, all = 0;
    for (int b = 0; b <= depth; ++b) {
        all += count_at(q, b);
        if (b > 0) path += count_at(q, b);
    }
    uint64_t* tot = reinterpret_cast<uint64_t*>(q + BQ_TOTALS64);
    tot[0] += count_at(q, 0);
    tot[1] += (uint64_t)S * all;
    tot[2] += path;
}

// --------------------------------------------------------------- primary
// raygenPrimary.comp:44-87 for every pixel of the views taking this pass
// (views are ordered by descending spp, so they form a prefix of the global
// pixel range); slot = global pixel.  Also resets the pixel's throughput.
__global__ __launch_bounds__(256) void batch_primary_kernel(const BatchTables t,
                                                            const uint32_t* __restrict__ q,
                                                            wf_ray* __restrict__ rays,
                                                            int32_t* __restrict__ s2i,
                                                            float4* __restrict__ dec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)count_at(q, 0)) return;
    const int v = view_of(t, i);
    const BatchView& V = t.views[v];
    const int local = (int)(i - t.offsets[v]);
    const int si = (int)q[BQ_PASS] + V.spp * ((int)q[BQ_FRAME] - 1);
    const V3 d = primary_direction(V.s2w, local, V.width, V.height, halton(si, 2), halton(si, 3),
                                   V.origin[0], V.origin[1], V.origin[2]);
    store_ray(rays, i, make_float4(V.origin[0], V.origin[1], V.origin[2], 0.0f),
              make_float4(d.x, d.y, d.z, V.max_dist));
    s2i[i] = (int32_t)i;
    dec[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
}

// ----------------------------------------------------------------- trace
// Ray i belongs to the path owner_s2i[i / mult] (mult = S for shadow rays).
template <bool CLOSEST, int BLOCK>
__global__ __launch_bounds__(BLOCK) void batch_trace_kernel(const BatchTables t,
                                                            const uint32_t* __restrict__ q, int b,
                                                            int mult,
                                                            const wf_ray* __restrict__ rays,
                                                            const int32_t* __restrict__ owner_s2i,
                                                            wf_hit* __restrict__ hits) {
    __shared__ int s_stack[LDS_DEPTH * BLOCK];
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= (int64_t)count_at(q, b) * mult) return;
    const int v = view_of(t, owner_s2i[i / mult]);
    const BatchScene& sc = t.scenes[t.views[v].scene];
    float4 ro, rd;
    load_ray(rays, i, ro, rd);
    WtRay wr;
    const bool live = (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
    const TraceOut tr =
        traverse<CLOSEST, BLOCK, false>(sc.nodes4, sc.tris, ro, rd, wr, live, s_stack);
    *reinterpret_cast<int4*>(hits + i) = hit_record(sc.tris, wr, tr);
}

// --------------------------------------------------------------- smooth
// reconstructSmooth.comp:34-97 (plus emission); the pixel's auxiliary value
// is written, not accumulated: each pixel has one ray per vertex and the
// single-view loop clears it before every vertex.
__global__ __launch_bounds__(256) void batch_smooth_kernel(
    const BatchTables t, const uint32_t* __restrict__ q, int b, const wf_ray* __restrict__ rays,
    const wf_hit* __restrict__ hits, const int32_t* __restrict__ s2i, float4* __restrict__ aux,
    const float4* __restrict__ dec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)count_at(q, b)) return;
    const int gp = s2i[i];
    const int v = view_of(t, gp);
    const BatchView& V = t.views[v];
    const float spp = (float)V.spp;
    const float4 dc = dec[gp];
    const int4 hit = *reinterpret_cast<const int4*>(hits + i);
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (hit.x == -1) {
        cr += dc.x * V.background[0] / spp;
        cg += dc.y * V.background[1] / spp;
        cb += dc.z * V.background[2] / spp;
    } else {
        const BatchScene& sc = t.scenes[V.scene];
        float4 ro, rd;
        load_ray(rays, i, ro, rd);
        const int64_t tri = (int64_t)sc.geom_first[hit.y] + hit.x;
        V3 nrm;
        const int m = clamp_material(shading_normal(sc.shade, tri, (uint32_t)hit.w, nrm), sc.nmat);
        const float th = __int_as_float(hit.z);
        const V3 p = v3(ro.x + rd.x * th, ro.y + rd.y * th, ro.z + rd.z * th);
        const V3 ff = faceforward3(nrm, v3(rd.x, rd.y, rd.z));
        const float alpha = smooth_alpha(ff, p, V.light[0], V.light[1], V.light[2]);
        cr += dc.x * (alpha + sc.emission[3 * m]) / spp;
        cg += dc.y * (alpha + sc.emission[3 * m + 1]) / spp;
        cb += dc.z * (alpha + sc.emission[3 * m + 2]) / spp;
    }
    aux[gp] = make_float4(cr, cg, cb, 1.0f);
}

// ---------------------------------------------------------- shadow rays
// raygenShadow.comp:68-139, hash keyed by the view's pixel index.
__global__ __launch_bounds__(256) void batch_shadow_gen_kernel(
    const BatchTables t, const uint32_t* __restrict__ q, int b, const uint32_t* __restrict__ seeds,
    const wf_ray* __restrict__ rays, const wf_hit* __restrict__ hits,
    const int32_t* __restrict__ s2i, wf_ray* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)count_at(q, b)) return;
    const int gp = s2i[j];
    const int v = view_of(t, gp);
    const BatchView& V = t.views[v];
    float4 ro, rd;
    load_ray(rays, j, ro, rd);
    const int4 hit = *reinterpret_cast<const int4*>(hits + j);
    const V3 off = shadow_offset(seeds[gp], (int)(gp - t.offsets[v]));
    const bool miss = hit.x == -1;
    const V3 o = shadow_origin(ro, rd, __int_as_float(hit.z), miss);
    const int S = t.shadow_samples;
    for (int k = 0; k < S; ++k) {
        float4 ra, rb;
        shadow_ray(k, S, off, o, miss, V.light[0], V.light[1], V.light[2], V.radius, ra, rb);
        store_ray(out, (int64_t)S * j + k, ra, rb);
    }
}

// reconstructShadow.comp:30-62: visible fraction of the S shadow rays
// times the auxiliary value, added to the frame.
__global__ __launch_bounds__(256) void batch_shadow_recon_kernel(
    const uint32_t* __restrict__ q, int b, int S, const wf_hit* __restrict__ shadow_hits,
    const int32_t* __restrict__ s2i, const float4* __restrict__ aux, float4* __restrict__ frame) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (int64_t)count_at(q, b)) return;
    const int gp = s2i[j];
    const float4 c = aux[gp];
    float shadow = 0.0f;
    for (int k = 0; k < S; ++k)
        if (shadow_hits[(int64_t)S * j + k].prim == -1) shadow += 1.0f;
    shadow /= (float)S;
    const float4 o = frame[gp];
    frame[gp] = make_float4(o.x + c.x * shadow, o.y + c.y * shadow, o.z + c.z * shadow, o.w + 1.0f);
}

// ----------------------------------------------------------- shade/extend
// raygenPath.comp:78-170 (no roulette) with order-preserving look-back
// compaction; the survivor count goes to q[count b + 1]. A workgroup takes
// BSHADE_TILE slots, BSHADE_ITEMS per thread (slot tile * BSHADE_TILE + item
// * BSHADE_BLOCK + thread), which keeps the look-back chain short.
constexpr int BSHADE_BLOCK = 256;
constexpr int BSHADE_ITEMS = 4;
constexpr int BSHADE_TILE = BSHADE_BLOCK * BSHADE_ITEMS;

__global__ __launch_bounds__(BSHADE_BLOCK) void batch_shade_kernel(
    const BatchTables t, uint32_t* __restrict__ q, int b, uint32_t* __restrict__ seeds,
    const wf_ray* __restrict__ in_rays, const wf_hit* __restrict__ hits,
    const int32_t* __restrict__ in_s2i, wf_ray* __restrict__ out_rays,
    int32_t* __restrict__ out_s2i, float4* __restrict__ dec, uint32_t* __restrict__ tile_counter,
    uint32_t* __restrict__ status) {
    __shared__ uint32_t s_tile, s_excl;
    __shared__ uint32_t s_wave[BSHADE_ITEMS][BSHADE_BLOCK / 64];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int64_t n = count_at(q, b);
    const int64_t ntiles = (n + BSHADE_TILE - 1) / BSHADE_TILE;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= ntiles) return;  // surplus workgroups of the capacity-sized grid

    bool alive[BSHADE_ITEMS];
    int gp[BSHADE_ITEMS];
    float4 oa[BSHADE_ITEMS], ob[BSHADE_ITEMS];
    float wr[BSHADE_ITEMS], wg[BSHADE_ITEMS], wb[BSHADE_ITEMS];
    uint32_t rank[BSHADE_ITEMS];
#pragma unroll
    for (int it = 0; it < BSHADE_ITEMS; ++it) {
        const int64_t j = tile * BSHADE_TILE + (int64_t)it * BSHADE_BLOCK + tid;
        alive[it] = false;
        gp[it] = 0;
        wr[it] = wg[it] = wb[it] = 1.f;
        if (j < n) {
            gp[it] = in_s2i[j];
            const int4 hit = *reinterpret_cast<const int4*>(hits + j);
            if (hit.x >= 0) {
                const int v = view_of(t, gp[it]);
                const BatchScene& sc = t.scenes[t.views[v].scene];
                uint32_t seed = seeds[gp[it]];
                float4 ro, rd;
                load_ray(in_rays, j, ro, rd);
                const int64_t tri = (int64_t)sc.geom_first[hit.y] + hit.x;
                V3 nrm;
                const int m = clamp_material(shading_normal(sc.shade, tri, (uint32_t)hit.w, nrm), sc.nmat);
                const V3 ff = faceforward3(nrm, v3(rd.x, rd.y, rd.z));
                const float r0 = randf(seed);
                const float r1 = randf(seed);
                const float th = __int_as_float(hit.z);
                const V3 p = v3(ro.x + rd.x * th, ro.y + rd.y * th, ro.z + rd.z * th);
                const float eps = 1.0e-4f;
                const V3 d = cosine_direction(r0, r1, ff);
                oa[it] = make_float4(p.x + d.x * eps, p.y + d.y * eps, p.z + d.z * eps, eps);
                ob[it] = make_float4(d.x, d.y, d.z, FLT_MAX);
                wr[it] = sc.albedo[3 * m];
                wg[it] = sc.albedo[3 * m + 1];
                wb[it] = sc.albedo[3 * m + 2];
                alive[it] = true;
                seeds[gp[it]] = seed;
            }
        }
        const uint64_t bal = ballot64(alive[it]);
        rank[it] = popc_below(bal);
        if (lane == 0) s_wave[it][wave] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
    uint32_t base[BSHADE_ITEMS], total = 0;
#pragma unroll
    for (int it = 0; it < BSHADE_ITEMS; ++it) {
#pragma unroll
        for (int w = 0; w < BSHADE_BLOCK / 64; ++w) {
            if (w == (int)wave) base[it] = total;
            total += s_wave[it][w];
        }
    }
    if (tid == 0) {
        uint32_t excl = 0;
        if (tile == 0) {
            lb_store(&status[0], LB_FLAG_INC | total);
        } else {
            lb_store(&status[tile], LB_FLAG_AGG | total);
            int64_t k = tile - 1;
            while (true) {
                uint32_t v;
                do {
                    v = lb_load(&status[k]);
                } while ((v & ~LB_MASK) == 0u);
                excl += v & LB_MASK;
                if (v & LB_FLAG_INC) break;
                --k;
            }
            lb_store(&status[tile], LB_FLAG_INC | (excl + total));
        }
        s_excl = excl;
        if (tile == ntiles - 1) q[BQ_COUNT0 + b + 1] = excl + total;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BSHADE_ITEMS; ++it) {
        if (alive[it]) {
            const int64_t k = (int64_t)s_excl + base[it] + rank[it];
            store_ray(out_rays, k, oa[it], ob[it]);
            out_s2i[k] = gp[it];
            const float4 dc = dec[gp[it]];
            dec[gp[it]] = make_float4(dc.x * wr[it], dc.y * wg[it], dc.z * wb[it], dc.w * 1.0f);
        }
    }
}

// ------------------------------------------------------ keys and reorder
// 64-bit key: 6-bit scene index, then the 30-bit origin Morton code in the
// scene's AABB and the 28-bit octahedral direction code, concatenated
// (MORTON5D = false) or interleaved as a 5-D Morton code (true: a wave then
// holds rays of a coarse origin cell and direction cone, as for the
// single-view 60-bit key).
template <bool MORTON5D>
__global__ __launch_bounds__(256) void batch_keys_kernel(const BatchTables t,
                                                         const uint32_t* __restrict__ q, int b,
                                                         const wf_ray* __restrict__ rays,
                                                         const int32_t* __restrict__ s2i,
                                                         uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)count_at(q, b)) return;
    const int sid = t.views[view_of(t, s2i[i])].scene;
    const BatchScene& sc = t.scenes[sid];
    const KeyFrame f{sc.bmin[0], sc.bmin[1], sc.bmin[2], sc.bsize[0], sc.bsize[1], sc.bsize[2]};
    float4 o, d;
    load_ray(rays, i, o, d);
    keys[i] = ((uint64_t)(sid & 63) << 58) |
              (MORTON5D ? morton5d<14>(f, o, d)
                        : (((uint64_t)origin30(f, o) << 28) | (uint64_t)oct_code(d, 14)));
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void batch_reorder_kernel(const uint32_t* __restrict__ q, int b,
                                                            const uint32_t* __restrict__ idx,
                                                            const wf_ray* __restrict__ in_rays,
                                                            const int32_t* __restrict__ in_s2i,
                                                            wf_ray* __restrict__ out_rays,
                                                            int32_t* __restrict__ out_s2i) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)count_at(q, b)) return;
    const uint32_t src = idx[i];
    float4 ra, rb;
    load_ray(in_rays, src, ra, rb);
    store_ray(out_rays, i, ra, rb);
    out_s2i[i] = in_s2i[src];
}

// ------------------------------------------------------------------ host
template <bool CLOSEST>
static void launch_batch_trace(const BatchTuning& tu, const BatchTables& t, const uint32_t* q, int b,
                               int mult, const wf_ray* rays, const int32_t* owner, wf_hit* hits,
                               int64_t cap, hipStream_t s) {
    switch (tu.block) {
        case 128:
            batch_trace_kernel<CLOSEST, 128><<<grid_for(cap, 128), 128, 0, s>>>(t, q, b, mult, rays, owner, hits);
            break;
        case 256:
            batch_trace_kernel<CLOSEST, 256><<<grid_for(cap, 256), 256, 0, s>>>(t, q, b, mult, rays, owner, hits);
            break;
        default:
            batch_trace_kernel<CLOSEST, 64><<<grid_for(cap, 64), 64, 0, s>>>(t, q, b, mult, rays, owner, hits);
            break;
    }
}

int batch_init_frame(const BatchTables& t, const BatchBuffers& bb, int64_t total_pixels, int frame,
                     hipStream_t s) {
    batch_seeds_kernel<<<grid_for(total_pixels, 256), 256, 0, s>>>(t, bb.seeds, total_pixels,
                                                                    (uint32_t)frame);
    WF_TRY(launch_check("batch_seeds_kernel"));
    WF_TRY(fill_u32(reinterpret_cast<uint32_t*>(bb.frame), 4 * total_pixels, 0u, s));
    WF_TRY(fill_u32(bb.q, BQ_WORDS, 0u, s));
    return WF_OK;
}

int batch_set_pass(const BatchBuffers& bb, int frame, int pass, uint32_t count, int depth,
                   hipStream_t s) {
    batch_set_pass_kernel<<<1, 64, 0, s>>>(bb.q, frame, pass, count, depth);
    return launch_check("batch_set_pass_kernel");
}

// One vertex: closest hits of the current rays, smooth term, shadow rays,
// their any-hit trace and the shadow reconstruction.
static int batch_vertex(const BatchTables& t, const BatchBuffers& bb, const BatchTuning& tu, int b,
                        int cur, hipStream_t s) {
    const int64_t cap = bb.capacity;
    const int S = t.shadow_samples;
    launch_batch_trace<true>(tu, t, bb.q, b, 1, bb.rays[cur], bb.s2i[cur], bb.hits, cap, s);
    WF_TRY(launch_check("batch_trace_kernel<closest>"));
    batch_smooth_kernel<<<grid_for(cap, 256), 256, 0, s>>>(t, bb.q, b, bb.rays[cur], bb.hits,
                                                          bb.s2i[cur], bb.aux, bb.dec);
    WF_TRY(launch_check("batch_smooth_kernel"));
    batch_shadow_gen_kernel<<<grid_for(cap, 256), 256, 0, s>>>(t, bb.q, b, bb.seeds, bb.rays[cur],
                                                              bb.hits, bb.s2i[cur], bb.shadow_rays);
    WF_TRY(launch_check("batch_shadow_gen_kernel"));
    launch_batch_trace<false>(tu, t, bb.q, b, S, bb.shadow_rays, bb.s2i[cur], bb.shadow_hits,
                              cap * S, s);
    WF_TRY(launch_check("batch_trace_kernel<any>"));
    batch_shadow_recon_kernel<<<grid_for(cap, 256), 256, 0, s>>>(bb.q, b, S, bb.shadow_hits,
                                                                bb.s2i[cur], bb.aux, bb.frame);
    return launch_check("batch_shadow_recon_kernel");
}

int batch_enqueue_pass(const BatchTables& t, const BatchBuffers& bb, const BatchTuning& tu,
                       int depth, hipStream_t s) {
    const int64_t cap = bb.capacity;
    int cur = 0;
    batch_primary_kernel<<<grid_for(cap, 256), 256, 0, s>>>(t, bb.q, bb.rays[0], bb.s2i[0], bb.dec);
    WF_TRY(launch_check("batch_primary_kernel"));
    WF_TRY(batch_vertex(t, bb, tu, 0, cur, s));
    const int64_t tiles = (cap + BSHADE_TILE - 1) / BSHADE_TILE;
    for (int b = 0; b < depth; ++b) {
        WF_TRY(fill_u32(bb.compact, 64 + tiles, 0u, s));
        batch_shade_kernel<<<(unsigned)tiles, BSHADE_BLOCK, 0, s>>>(
            t, bb.q, b, bb.seeds, bb.rays[cur], bb.hits, bb.s2i[cur], bb.rays[1 - cur],
            bb.s2i[1 - cur], bb.dec, bb.compact, bb.compact + 64);
        WF_TRY(launch_check("batch_shade_kernel"));
        cur = 1 - cur;
        if (tu.reorder_period > 0 && b % tu.reorder_period == 0) {
            if (tuning().batch_key_layout == 1)
                batch_keys_kernel<true><<<grid_for(cap, 256), 256, 0, s>>>(t, bb.q, b + 1, bb.rays[cur],
                                                                           bb.s2i[cur], bb.keys, bb.idx);
            else
                batch_keys_kernel<false><<<grid_for(cap, 256), 256, 0, s>>>(t, bb.q, b + 1, bb.rays[cur],
                                                                            bb.s2i[cur], bb.keys, bb.idx);
            WF_TRY(launch_check("batch_keys_kernel"));
            WF_TRY(sort_pairs_launch(bb.keys, bb.idx, cap, bb.q + BQ_COUNT0 + b + 1, 8, 64,
                                     tu.digit_bits, bb.sort_scratch, bb.sort_scratch_bytes, s));
            batch_reorder_kernel<<<grid_for(cap, 256), 256, 0, s>>>(
                bb.q, b + 1, bb.idx, bb.rays[cur], bb.s2i[cur], bb.rays[1 - cur], bb.s2i[1 - cur]);
            WF_TRY(launch_check("batch_reorder_kernel"));
            cur = 1 - cur;
        }
        WF_TRY(batch_vertex(t, bb, tu, b + 1, cur, s));
    }
    batch_totals_kernel<<<1, 64, 0, s>>>(bb.q, depth, t.shadow_samples);
    return launch_check("batch_totals_kernel");
}

}  // namespace wfpt
