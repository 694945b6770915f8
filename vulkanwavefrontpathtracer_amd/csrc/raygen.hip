// Ray generation stages (RayGen, projects/WaveFrontPathTracer/Ray/RayGen.cpp)
// and shade/extend with deterministic stream compaction.
#include "shading.h"

#include <cfloat>

namespace wfpt {

// ------------------------------------------------------------------ seeds
// initSeeds.comp:31-42.
__global__ __launch_bounds__(256) void init_seeds_kernel(uint32_t* __restrict__ seeds, int64_t n,
                                                         uint32_t frame) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seeds[i] = tea((uint32_t)i, frame, 16);
}

// ---------------------------------------------------------------- primary
// raygenPrimary.comp:44-87.
__global__ __launch_bounds__(256) void raygen_primary_kernel(const Mat4 s2w, float ox, float oy,
                                                             float oz,
                                                             const int32_t* __restrict__ i2p,
                                                             wf_ray* __restrict__ rays, int width,
                                                             int height, float hx, float hy,
                                                             float max_dist) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t np = (int64_t)width * height;
    if (i >= np) return;
    V3 d = primary_direction(s2w, i2p[i], width, height, hx, hy, ox, oy, oz);
    store_ray(rays, i, make_float4(ox, oy, oz, 0.0f), make_float4(d.x, d.y, d.z, max_dist));
}

// ----------------------------------------------------------------- shadow
// raygenShadow.comp:68-139; the seed and the hash are keyed by the path's
// pixel (in_s2i) instead of the slot.
__global__ __launch_bounds__(256) void raygen_shadow_kernel(
    const uint32_t* __restrict__ seeds, const wf_ray* __restrict__ in_rays,
    const wf_hit* __restrict__ in_hits, const int32_t* __restrict__ in_s2i,
    wf_ray* __restrict__ out_rays, int32_t* __restrict__ out_s2i, int32_t* __restrict__ out_i2s,
    int64_t begin, int64_t count, int S, float lx, float ly, float lz, float radius) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    int64_t src = begin + j;
    int pixel = in_s2i[src];
    float4 ro, rd;
    load_ray(in_rays, src, ro, rd);
    const int4 hit = *reinterpret_cast<const int4*>(in_hits + src);
    const V3 off = shadow_offset(seeds[pixel], pixel);
    const bool miss = hit.x == -1;
    const V3 o = shadow_origin(ro, rd, __int_as_float(hit.z), miss);
    int64_t out = (int64_t)S * j;
    for (int i = 0; i < S; ++i, ++out) {
        float4 ra, rb;
        shadow_ray(i, S, off, o, miss, lx, ly, lz, radius, ra, rb);
        store_ray(out_rays, out, ra, rb);
        out_s2i[out] = (int32_t)out;
        out_i2s[out] = (int32_t)out;
    }
}

// ----------------------------------------------------------- shade/extend
constexpr int SHADE_BLOCK = 256;
constexpr int SHADE_ITEMS = 4;                       // slots per thread
constexpr int SHADE_TILE = SHADE_BLOCK * SHADE_ITEMS;

// raygenPath.comp:78-170. A workgroup takes a tile of SHADE_TILE input
// slots, one per thread and item (slot tile * SHADE_TILE + item *
// SHADE_BLOCK + thread). Survivors are written in input order: each
// workgroup counts its survivors, resolves the count of earlier workgroups
// by decoupled look-back (tiles taken in arrival order), then scatters.
// Several slots per thread keep the look-back chain (one status word per
// tile, published and read at agent scope) short against the shading work.
__global__ __launch_bounds__(SHADE_BLOCK) void shade_extend_kernel(
    const float4* __restrict__ shade, const int32_t* __restrict__ geom_first,
    const float* __restrict__ albedo, int nmat, uint32_t* __restrict__ seeds,
    const wf_ray* __restrict__ in_rays, const wf_hit* __restrict__ in_hits,
    const int32_t* __restrict__ in_s2i, int64_t n, wf_ray* __restrict__ out_rays,
    int32_t* __restrict__ out_s2i, float4* __restrict__ decreases, uint32_t* __restrict__ d_count,
    uint32_t* __restrict__ tile_counter, uint32_t* __restrict__ status, int rr, int64_t ntiles,
    uint8_t* __restrict__ sign_log) {
    __shared__ uint32_t s_tile, s_excl;
    __shared__ uint32_t s_wave[SHADE_ITEMS][SHADE_BLOCK / 64];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const int64_t tile = s_tile;

    bool alive[SHADE_ITEMS];
    int pixel[SHADE_ITEMS];
    float4 oa[SHADE_ITEMS], ob[SHADE_ITEMS];
    float wr[SHADE_ITEMS], wg[SHADE_ITEMS], wb[SHADE_ITEMS];
    uint32_t rank[SHADE_ITEMS];
#pragma unroll
    for (int it = 0; it < SHADE_ITEMS; ++it) {
        const int64_t j = tile * SHADE_TILE + (int64_t)it * SHADE_BLOCK + tid;
        alive[it] = false;
        pixel[it] = 0;
        wr[it] = wg[it] = wb[it] = 1.f;
        if (j < n) {
            pixel[it] = in_s2i[j];
            const int4 hit = *reinterpret_cast<const int4*>(in_hits + j);
            if (hit.x >= 0) {
                uint32_t seed = seeds[pixel[it]];
                float4 ro, rd;
                load_ray(in_rays, j, ro, rd);
                int64_t tri = (int64_t)geom_first[hit.y] + hit.x;
                V3 nrm;
                int m = clamp_material(shading_normal(shade, tri, (uint32_t)hit.w, nrm), nmat);
                V3 ff = faceforward3(nrm, v3(rd.x, rd.y, rd.z));
                // decision log: the normal flipped to face the ray (bit 0) and the
                // basis branch of a normal with negative z (bit 1)
                if (sign_log)
                    sign_log[pixel[it]] = (uint8_t)((dot3(nrm, v3(rd.x, rd.y, rd.z)) < 0.0f ? 0 : 1) |
                                                    (ff.z >= 0.0f ? 0 : 2));
                float r0 = randf(seed);
                float r1 = randf(seed);
                float t = __int_as_float(hit.z);
                V3 p = v3(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
                const float eps = 1.0e-4f;
                V3 d = cosine_direction(r0, r1, ff);
                oa[it] = make_float4(p.x + d.x * eps, p.y + d.y * eps, p.z + d.z * eps, eps);
                ob[it] = make_float4(d.x, d.y, d.z, FLT_MAX);
                wr[it] = albedo[3 * m];
                wg[it] = albedo[3 * m + 1];
                wb[it] = albedo[3 * m + 2];
                alive[it] = true;
                if (rr) {
                    // Russian roulette on the throughput after this bounce.
                    float4 dec = decreases[pixel[it]];
                    float q = fmaxf(fmaxf(dec.x * wr[it], dec.y * wg[it]), dec.z * wb[it]);
                    q = fminf(fmaxf(q, 0.05f), 1.0f);
                    float r2 = randf(seed);
                    if (r2 >= q) {
                        alive[it] = false;
                    } else {
                        wr[it] /= q;
                        wg[it] /= q;
                        wb[it] /= q;
                    }
                }
                seeds[pixel[it]] = seed;
            }
        }
        // rank among the item's survivors of this wave
        const uint64_t bal = ballot64(alive[it]);
        rank[it] = popc_below(bal);
        if (lane == 0) s_wave[it][wave] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
    // survivors before (item, wave) in slot order, and the tile's total
    uint32_t base[SHADE_ITEMS], total = 0;
#pragma unroll
    for (int it = 0; it < SHADE_ITEMS; ++it) {
#pragma unroll
        for (int w = 0; w < SHADE_BLOCK / 64; ++w) {
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
        if (tile == ntiles - 1) *d_count = excl + total;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SHADE_ITEMS; ++it) {
        if (alive[it]) {
            const int64_t k = (int64_t)s_excl + base[it] + rank[it];
            store_ray(out_rays, k, oa[it], ob[it]);
            out_s2i[k] = pixel[it];
            float4 dec = decreases[pixel[it]];
            decreases[pixel[it]] = make_float4(dec.x * wr[it], dec.y * wg[it], dec.z * wb[it], dec.w * 1.0f);
        }
    }
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_init_seeds(uint32_t* seeds, int64_t n, int frame_index, wf_stream_t stream,
                             float* ms) {
    WF_CHECK_ARG(n >= 0, "wf_init_seeds: n < 0");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(seeds, "wf_init_seeds: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    init_seeds_kernel<<<grid_for(n, 256), 256, 0, s>>>(seeds, n, (uint32_t)frame_index);
    WF_TRY(launch_check("init_seeds_kernel"));
    return t.finish();
}

// raygenPrimary.comp:11-26: halton2D(sampleIndex), identical for all pixels;
// evaluated on the host with the same float operations.
static float halton_host(int i, int base) {
    float r = 0.0f;
    float f = 1.0f / (float)base;
    int j = i;
    while (j > 0) {
        r += f * (float)(j % base);
        j /= base;
        f /= (float)base;
    }
    return r;
}

extern "C" int wf_raygen_primary(const float screen_to_world[16], const float origin[3],
                                 const int32_t* index_to_pixel, wf_ray* rays, int width, int height,
                                 int sample_index, float max_dist, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(screen_to_world && origin, "wf_raygen_primary: null matrix/origin");
    WF_CHECK_ARG(width > 0 && height > 0, "wf_raygen_primary: empty image");
    WF_CHECK_ARG(index_to_pixel && rays, "wf_raygen_primary: null buffer");
    Mat4 m;
    for (int k = 0; k < 16; ++k) m.m[k] = screen_to_world[k];
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    int64_t np = (int64_t)width * height;
    raygen_primary_kernel<<<grid_for(np, 256), 256, 0, s>>>(
        m, origin[0], origin[1], origin[2], index_to_pixel, rays, width, height,
        halton_host(sample_index, 2), halton_host(sample_index, 3), max_dist);
    WF_TRY(launch_check("raygen_primary_kernel"));
    return t.finish();
}

extern "C" int wf_raygen_shadow(const uint32_t* seeds, const wf_ray* in_rays, const wf_hit* in_hits,
                                const int32_t* in_s2i, wf_ray* out_rays, int32_t* out_s2i,
                                int32_t* out_i2s, int64_t batch_begin, int64_t batch_end,
                                int num_samples, const float light[3], float light_radius,
                                wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(batch_begin >= 0 && batch_end >= batch_begin, "wf_raygen_shadow: bad batch");
    WF_CHECK_ARG(num_samples >= 1 && light, "wf_raygen_shadow: need >= 1 sample and a light");
    int64_t count = batch_end - batch_begin;
    if (count == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(seeds && in_rays && in_hits && in_s2i && out_rays && out_s2i && out_i2s,
                 "wf_raygen_shadow: null buffer");
    WF_CHECK_ARG(count * num_samples <= (int64_t)INT32_MAX, "wf_raygen_shadow: batch too large");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    raygen_shadow_kernel<<<grid_for(count, 256), 256, 0, s>>>(
        seeds, in_rays, in_hits, in_s2i, out_rays, out_s2i, out_i2s, batch_begin, count,
        num_samples, light[0], light[1], light[2], light_radius);
    WF_TRY(launch_check("raygen_shadow_kernel"));
    return t.finish();
}

extern "C" int64_t wf_compact_scratch_bytes(int64_t n) {
    if (n < 0) return WF_ERR_INVALID_ARG;
    int64_t tiles = (n + SHADE_TILE - 1) / SHADE_TILE;
    return 256 + 4 * (tiles + 1);
}

namespace wfpt {
int shade_extend_logged(wf_scene scene, uint32_t* seeds, const wf_ray* in_rays,
                        const wf_hit* in_hits, const int32_t* in_s2i, int64_t n, wf_ray* out_rays,
                        int32_t* out_s2i, float* decreases, uint32_t* d_count,
                        int russian_roulette, int bounce, int rr_start, void* scratch,
                        int64_t scratch_bytes, uint32_t* host_count, wf_stream_t stream, float* ms,
                        uint8_t* sign_log);
}  // namespace wfpt

extern "C" int wf_shade_extend(wf_scene scene, uint32_t* seeds, const wf_ray* in_rays,
                               const wf_hit* in_hits, const int32_t* in_s2i, int64_t n,
                               wf_ray* out_rays, int32_t* out_s2i, float* decreases,
                               uint32_t* d_count, int russian_roulette, int bounce, int rr_start,
                               void* scratch, int64_t scratch_bytes, uint32_t* host_count,
                               wf_stream_t stream, float* ms) {
    return wfpt::shade_extend_logged(scene, seeds, in_rays, in_hits, in_s2i, n, out_rays, out_s2i,
                                     decreases, d_count, russian_roulette, bounce, rr_start, scratch,
                                     scratch_bytes, host_count, stream, ms, nullptr);
}

// wf_shade_extend with the frame's decision log of the bounce signs
// (sign_log[pixel], or nullptr).
int wfpt::shade_extend_logged(wf_scene scene, uint32_t* seeds, const wf_ray* in_rays,
                              const wf_hit* in_hits, const int32_t* in_s2i, int64_t n,
                              wf_ray* out_rays, int32_t* out_s2i, float* decreases,
                              uint32_t* d_count, int russian_roulette, int bounce, int rr_start,
                              void* scratch, int64_t scratch_bytes, uint32_t* host_count,
                              wf_stream_t stream, float* ms, uint8_t* sign_log) {
    WF_CHECK_ARG(scene, "wf_shade_extend: null scene");
    WF_CHECK_ARG(n >= 0, "wf_shade_extend: n < 0");
    WF_CHECK_ARG(d_count, "wf_shade_extend: null d_count");
    if (n > WF_MAX_ELEMENTS) return fail(WF_ERR_TOO_LARGE, "wf_shade_extend: n exceeds 2^30-1");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    if (n == 0) {
        WF_TRY(fill_u32(d_count, 1, 0u, s));
    } else {
        WF_CHECK_ARG(seeds && in_rays && in_hits && in_s2i && out_rays && out_s2i && decreases &&
                         scratch,
                     "wf_shade_extend: null buffer");
        if (scratch_bytes < wf_compact_scratch_bytes(n))
            return fail(WF_ERR_SCRATCH_TOO_SMALL, "wf_shade_extend: scratch < wf_compact_scratch_bytes");
        int64_t tiles = (n + SHADE_TILE - 1) / SHADE_TILE;
        uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch);
        uint32_t* status = ctr + 64;
        WF_TRY(fill_u32(ctr, 64 + tiles, 0u, s));
        int rr = (russian_roulette && bounce >= rr_start) ? 1 : 0;
        shade_extend_kernel<<<(unsigned)tiles, SHADE_BLOCK, 0, s>>>(
            scene->shade, scene->geom_first, scene->albedo, scene->num_materials, seeds, in_rays,
            in_hits, in_s2i, n, out_rays, out_s2i, reinterpret_cast<float4*>(decreases), d_count,
            ctr, status, rr, tiles, sign_log);
        WF_TRY(launch_check("shade_extend_kernel"));
    }
    if (host_count) {
        WF_HIP(hipMemcpyAsync(host_count, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        WF_HIP(hipStreamSynchronize(s));
    }
    return t.finish();
}
