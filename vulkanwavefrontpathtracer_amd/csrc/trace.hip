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

namespace wfpt {

template <bool CLOSEST, int BLOCK, bool LDS, bool STATS = false>
__global__ __launch_bounds__(BLOCK) void trace_kernel(const float4* __restrict__ nodes,
                                                      const float4* __restrict__ tris,
                                                      const wf_ray* __restrict__ rays,
                                                      const uint32_t* __restrict__ ray_index,
                                                      wf_hit* __restrict__ hits, int64_t n,
                                                      uint32_t* __restrict__ stat_nodes = nullptr,
                                                      uint32_t* __restrict__ stat_tris = nullptr) {
    __shared__ int s_stack[LDS ? LDS_DEPTH * BLOCK : 1];
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int64_t r = ray_index ? (int64_t)ray_index[i] : i;
    float4 ro, rd;
    load_ray(rays, r, ro, rd);
    WtRay wr;
    const bool live = (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
    const TraceOut tr = traverse<CLOSEST, BLOCK, LDS, STATS>(nodes, tris, ro, rd, wr, live, s_stack);
    *reinterpret_cast<int4*>(hits + r) = hit_record(tris, wr, tr);
    if (STATS) {
        stat_nodes[r] = tr.nodes_visited;
        stat_tris[r] = tr.tris_tested;
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
