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

template <bool CLOSEST, int BLOCK, bool LDS, bool STATS = false>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const float4* __restrict__ nodes,
                                                      const float4* __restrict__ tris,
                                                      const wf_ray* __restrict__ rays,
                                                      const uint32_t* __restrict__ ray_index,
                                                      wf_hit* __restrict__ hits, int64_t n,
                                                      uint32_t* __restrict__ stat_nodes = nullptr,
                                                      uint32_t* __restrict__ stat_tris = nullptr) {
    uint32_t n_nodes = 0, n_tris = 0;
    __shared__ int s_stack[LDS ? LDS_DEPTH * BLOCK : 1];
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int64_t r = ray_index ? (int64_t)ray_index[i] : i;
    float4 ro, rd;
    load_ray(rays, r, ro, rd);
    const float tmin = ro.w, tmax = rd.w;

    WtRay wr;
    bool live = (tmax >= tmin) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
    int best_slot = -1;
    int best_orig = 0x7fffffff;
    float best_t = tmax;
    if (live) {
        const float ix = safe_inv(rd.x), iy = safe_inv(rd.y), iz = safe_inv(rd.z);
        const float oix = ro.x * ix, oiy = ro.y * iy, oiz = ro.z * iz;
        int spill[LDS ? TRACE_STACK - LDS_DEPTH : TRACE_STACK];
        int sp = 0;
        auto push = [&](int v) {
            if (LDS && sp < LDS_DEPTH) s_stack[sp * BLOCK + threadIdx.x] = v;
            else spill[sp - (LDS ? LDS_DEPTH : 0)] = v;
            ++sp;
        };
        auto pop = [&]() {
            --sp;
            return (LDS && sp < LDS_DEPTH) ? s_stack[sp * BLOCK + threadIdx.x]
                                           : spill[sp - (LDS ? LDS_DEPTH : 0)];
        };
        int node = 0;
        while (true) {
            if (node >= 0) {
                if (STATS) ++n_nodes;
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
                    if (STATS) ++n_tris;
                    WtHit h;
                    if (!wt_intersect(wr, v0, v1, v2, h)) continue;
                    if (!(h.t >= tmin && h.t <= best_t)) continue;
                    const int orig = __float_as_int(v0.w);
                    if (CLOSEST) {
                        if (h.t < best_t || orig < best_orig) {
                            best_t = h.t;
                            best_orig = orig;
                            best_slot = first + k;
                        }
                    } else {
                        best_t = h.t;
                        best_orig = orig;
                        best_slot = first + k;
                        done = true;
                        break;
                    }
                }
                if (done || sp == 0) break;
                node = pop();
            }
        }
    }
    wf_hit out;
    if (best_slot >= 0) {
        const float4* tp = tris + 3 * (int64_t)best_slot;
        const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
        WtHit h;
        wt_intersect(wr, v0, v1, v2, h);
        out.prim = __float_as_int(v1.w);
        out.geom = __float_as_int(v2.w);
        out.t = best_t;
        out.bary = pack_unorm2x16(h.v / h.det, h.w / h.det);
    } else {
        out.prim = -1;
        out.geom = -1;
        out.t = -1.0f;
        out.bary = 0u;
    }
    *reinterpret_cast<int4*>(hits + r) = make_int4(out.prim, out.geom, __float_as_int(out.t), (int)out.bary);
    if (STATS) {
        stat_nodes[r] = n_nodes;
        stat_tris[r] = n_tris;
    }
}

__global__ __launch_bounds__(256) void miss_kernel(const uint32_t* __restrict__ ray_index,
                                                   wf_hit* __restrict__ hits, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t r = ray_index ? (int64_t)ray_index[i] : i;
    *reinterpret_cast<int4*>(hits + r) = make_int4(-1, -1, __float_as_int(-1.0f), 0);
}

template <bool CLOSEST, int BLOCK>
static void launch_trace(bool lds, const float4* nodes, const float4* tris, const wf_ray* rays,
                         const uint32_t* ray_index, wf_hit* hits, int64_t n, hipStream_t s) {
    unsigned g = grid_for(n, BLOCK);
    if (lds)
        trace_kernel<CLOSEST, BLOCK, true><<<g, BLOCK, 0, s>>>(nodes, tris, rays, ray_index, hits, n);
    else
        trace_kernel<CLOSEST, BLOCK, false><<<g, BLOCK, 0, s>>>(nodes, tris, rays, ray_index, hits, n);
}

template <bool CLOSEST>
static void launch_trace_block(int block, bool lds, const float4* nodes, const float4* tris,
                               const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                               int64_t n, hipStream_t s) {
    switch (block) {
        case 64: launch_trace<CLOSEST, 64>(lds, nodes, tris, rays, ray_index, hits, n, s); break;
        case 128: launch_trace<CLOSEST, 128>(lds, nodes, tris, rays, ray_index, hits, n, s); break;
        default: launch_trace<CLOSEST, 256>(lds, nodes, tris, rays, ray_index, hits, n, s); break;
    }
}

int trace_launch(wf_scene scene, const wf_ray* rays, const uint32_t* ray_index, wf_hit* hits,
                 int64_t n, int closest, hipStream_t s) {
    if (n == 0) return WF_OK;
    if (scene->num_tris == 0) {
        miss_kernel<<<grid_for(n, 256), 256, 0, s>>>(ray_index, hits, n);
        return launch_check("miss_kernel");
    }
    const Tuning& tu = tuning();
    if (closest)
        launch_trace_block<true>(tu.trace_block, tu.trace_lds_stack != 0, scene->nodes, scene->tris,
                                 rays, ray_index, hits, n, s);
    else
        launch_trace_block<false>(tu.trace_block, tu.trace_lds_stack != 0, scene->nodes, scene->tris,
                                  rays, ray_index, hits, n, s);
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
        trace_kernel<true, 64, true, true><<<grid_for(n, 64), 64, 0, s>>>(
            scene->nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
    else
        trace_kernel<false, 64, true, true><<<grid_for(n, 64), 64, 0, s>>>(
            scene->nodes, scene->tris, rays, nullptr, hits, n, node_visits, tri_tests);
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
