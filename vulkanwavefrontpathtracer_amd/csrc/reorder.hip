// Reorder gather through the sorted permutation.
//
// Restates reorderRays.comp:27-47 as driven by Tracer::reorderRays
// (projects/WaveFrontPathTracer/Tracer/Tracer.cpp:80-98):
//   out_rays[i] = in_rays[idx[i]]; out_s2i[i] = in_s2i[idx[i]];
//   out_i2s[out_s2i[i]] = i.
// Each thread moves one 32-byte ray as two 16-byte loads and stores; the
// index read and the ray/LUT writes are coalesced, the ray and LUT reads are
// gathers.
#include "common.h"

namespace wfpt {

__global__ __launch_bounds__(256) void reorder_kernel(const uint32_t* __restrict__ idx,
                                                      const wf_ray* __restrict__ in_rays,
                                                      wf_ray* __restrict__ out_rays,
                                                      const int32_t* __restrict__ in_s2i,
                                                      int32_t* __restrict__ out_s2i,
                                                      int32_t* __restrict__ out_i2s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t src = idx[i];
    float4 a, b;
    load_ray(in_rays, src, a, b);
    int32_t index = in_s2i[src];
    store_ray(out_rays, i, a, b);
    out_s2i[i] = index;
    out_i2s[index] = (int32_t)i;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_reorder(const uint32_t* idx, const wf_ray* in_rays, wf_ray* out_rays,
                          const int32_t* in_s2i, int32_t* out_s2i, int32_t* out_i2s, int64_t n,
                          wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(n >= 0, "wf_reorder: n < 0");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(idx && in_rays && out_rays && in_s2i && out_s2i && out_i2s,
                 "wf_reorder: null buffer");
    WF_CHECK_ARG((const void*)in_rays != (const void*)out_rays, "wf_reorder: in_rays == out_rays");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    reorder_kernel<<<grid_for(n, 256), 256, 0, s>>>(idx, in_rays, out_rays, in_s2i, out_s2i,
                                                    out_i2s, n);
    WF_TRY(launch_check("reorder_kernel"));
    return t.finish();
}
