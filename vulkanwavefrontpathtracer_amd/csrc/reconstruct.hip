// Film reconstruction stages of the Renderer
// (projects/WaveFrontPathTracer/Renderer/Renderer.cpp:201-245, :106-122,
// :54-91): smooth shading, shadow visibility, progressive accumulation with
// Reinhard tone mapping, hit counting.
#include "shading.h"

namespace wfpt {

// reconstructSmooth.comp:34-97.  The emission term is this library's
// addition (zero by default, which reproduces the reference).
__global__ __launch_bounds__(256) void reconstruct_smooth_kernel(
    const float4* __restrict__ shade, const int32_t* __restrict__ geom_first,
    const float* __restrict__ emission, int nmat, const wf_ray* __restrict__ rays,
    const wf_hit* __restrict__ hits, const int32_t* __restrict__ s2i, float4* __restrict__ pixels,
    const float4* __restrict__ decreases, int64_t n, float lx, float ly, float lz, float bgr,
    float bgg, float bgb, float spp) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int pixel = s2i[i];
    float4 col = pixels[pixel];
    float4 dec = decreases[pixel];
    const int4 hit = *reinterpret_cast<const int4*>(hits + i);
    float cr = col.x, cg = col.y, cb = col.z;
    if (hit.x == -1) {
        cr += dec.x * bgr / spp;
        cg += dec.y * bgg / spp;
        cb += dec.z * bgb / spp;
    } else {
        float4 ro, rd;
        load_ray(rays, i, ro, rd);
        int64_t tri = (int64_t)geom_first[hit.y] + hit.x;
        V3 nrm;
        int m = clamp_material(shading_normal(shade, tri, (uint32_t)hit.w, nrm), nmat);
        float t = __int_as_float(hit.z);
        V3 p = v3(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
        V3 ff = faceforward3(nrm, v3(rd.x, rd.y, rd.z));
        float alpha = smooth_alpha(ff, p, lx, ly, lz);
        float er = emission[3 * m], eg = emission[3 * m + 1], eb = emission[3 * m + 2];
        cr += dec.x * (alpha + er) / spp;
        cg += dec.y * (alpha + eg) / spp;
        cb += dec.z * (alpha + eb) / spp;
    }
    pixels[pixel] = make_float4(cr, cg, cb, 1.0f);
}

// reconstructShadow.comp:30-62.
__global__ __launch_bounds__(256) void reconstruct_shadow_kernel(
    const wf_hit* __restrict__ shadow_hits, const int32_t* __restrict__ in_s2i,
    const int32_t* __restrict__ shadow_i2s, const float4* __restrict__ in_pixels,
    float4* __restrict__ out_pixels, int64_t begin, int64_t count, int S, int replace) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    int pixel = in_s2i[begin + j];
    float4 c = in_pixels[pixel];
    float shadow = 0.0f;
    for (int i = 0; i < S; ++i) {
        int slot = shadow_i2s[j * S + i];
        if (shadow_hits[slot].prim == -1) shadow += 1.0f;
    }
    shadow /= (float)S;
    float4 v = make_float4(c.x * shadow, c.y * shadow, c.z * shadow, 1.0f);
    if (replace) {
        out_pixels[pixel] = v;
    } else {
        float4 o = out_pixels[pixel];
        out_pixels[pixel] = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
    }
}

// interpolateColors.comp:10-57.  GLSL's mat3(a..i) constructor fills columns,
// so `M * v` computes x = a*v.x + d*v.y + g*v.z: the rows of the effective
// matrix are the columns written in the shader.
__device__ __forceinline__ V3 mat3_mul_cols(const float* c, V3 v) {
    return v3((c[0] * v.x + c[3] * v.y) + c[6] * v.z, (c[1] * v.x + c[4] * v.y) + c[7] * v.z,
              (c[2] * v.x + c[5] * v.y) + c[8] * v.z);
}

__device__ __forceinline__ V3 tonemap_reinhard(V3 color, float key, float white) {
    const float rgb2xyz[9] = {0.5141364f, 0.3238786f, 0.16036376f, 0.265068f, 0.67023428f,
                              0.06409157f, 0.0241188f, 0.1228178f, 0.84442666f};
    const float xyz2rgb[9] = {2.5651f, -1.1665f, -0.3986f, -1.0217f, 1.9777f,
                              0.0439f, 0.0753f, -0.2543f, 1.1892f};
    const float avg_lum = 0.5f;
    V3 XYZ = mat3_mul_cols(rgb2xyz, color);
    float sum = (XYZ.x + XYZ.y) + XYZ.z;
    float Y = XYZ.y;
    float x = XYZ.x / sum;
    float y = XYZ.y / sum;
    float Lp = Y * key / avg_lum;
    Y = (Lp * (1.0f + Lp / (white * white))) / (1.0f + Lp);
    XYZ.x = Y * x / y;
    XYZ.y = Y;
    XYZ.z = Y * (1.0f - x - y) / y;
    return mat3_mul_cols(xyz2rgb, XYZ);
}

// interpolateColors.comp:71-100.
__global__ __launch_bounds__(256) void interpolate_kernel(float4* __restrict__ frame,
                                                          float4* __restrict__ pixels, int64_t n,
                                                          int frame_index, float key, float white) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (frame_index > 1) {
        float4 c0 = pixels[i], c1 = frame[i];
        float a = 1.0f / (float)frame_index;
        float b = a * ((float)frame_index - 1.0f);
        V3 c = v3(a * c1.x + b * c0.x, a * c1.y + b * c0.y, a * c1.z + b * c0.z);
        pixels[i] = make_float4(c.x, c.y, c.z, 1.0f);
        V3 t = tonemap_reinhard(c, key, white);
        frame[i] = make_float4(t.x, t.y, t.z, 1.0f);
    } else {
        float4 c = frame[i];
        pixels[i] = c;
        V3 t = tonemap_reinhard(v3(c.x, c.y, c.z), key, white);
        frame[i] = make_float4(t.x, t.y, t.z, 1.0f);
    }
}

// countRayHits.comp:22-51: wave sum -> LDS -> one atomic per workgroup. A
// workgroup takes COUNT_HITS_ITEMS x 256 consecutive records per round (the
// loads of a round all in flight together) and strides over the records
// with at most COUNT_HITS_GRID workgroups, so the one counter takes a few
// hundred atomics, which serialise on its address.
constexpr int COUNT_HITS_GRID = 512;
constexpr int COUNT_HITS_ITEMS = 8;
__global__ __launch_bounds__(256) void count_hits_kernel(const wf_hit* __restrict__ hits, int64_t n,
                                                         uint32_t* __restrict__ count) {
    __shared__ uint32_t s[4];
    uint32_t c = 0;
    const int64_t round = (int64_t)COUNT_HITS_ITEMS * 256;
    for (int64_t base = (int64_t)blockIdx.x * round; base < n; base += (int64_t)gridDim.x * round) {
        int prim[COUNT_HITS_ITEMS];
#pragma unroll
        for (int k = 0; k < COUNT_HITS_ITEMS; ++k) {
            const int64_t i = base + (int64_t)k * 256 + threadIdx.x;
            prim[k] = i < n ? hits[i].prim : -1;
        }
#pragma unroll
        for (int k = 0; k < COUNT_HITS_ITEMS; ++k) c += prim[k] >= 0 ? 1u : 0u;
    }
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane_id() == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = s[0] + s[1] + s[2] + s[3];
        if (t) atomicAdd(count, t);
    }
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_reconstruct_smooth(wf_scene scene, const wf_ray* rays, const wf_hit* hits,
                                     const int32_t* s2i, float* pixels, const float* decreases,
                                     int64_t n, const float light[3], const float background[3],
                                     int num_samples, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(scene && light && background, "wf_reconstruct_smooth: null scene/light/background");
    WF_CHECK_ARG(n >= 0 && num_samples >= 1, "wf_reconstruct_smooth: bad n or num_samples");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(rays && hits && s2i && pixels && decreases, "wf_reconstruct_smooth: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    reconstruct_smooth_kernel<<<grid_for(n, 256), 256, 0, s>>>(
        scene->shade, scene->geom_first, scene->emission, scene->num_materials, rays, hits, s2i,
        reinterpret_cast<float4*>(pixels), reinterpret_cast<const float4*>(decreases), n, light[0],
        light[1], light[2], background[0], background[1], background[2], (float)num_samples);
    WF_TRY(launch_check("reconstruct_smooth_kernel"));
    return t.finish();
}

extern "C" int wf_reconstruct_shadow(const wf_hit* shadow_hits, const int32_t* in_s2i,
                                     const int32_t* shadow_i2s, const float* in_pixels,
                                     float* out_pixels, int64_t batch_begin, int64_t batch_end,
                                     int num_samples, int replace, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(batch_begin >= 0 && batch_end >= batch_begin && num_samples >= 1,
                 "wf_reconstruct_shadow: bad batch / num_samples");
    int64_t count = batch_end - batch_begin;
    if (count == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(shadow_hits && in_s2i && shadow_i2s && in_pixels && out_pixels,
                 "wf_reconstruct_shadow: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    reconstruct_shadow_kernel<<<grid_for(count, 256), 256, 0, s>>>(
        shadow_hits, in_s2i, shadow_i2s, reinterpret_cast<const float4*>(in_pixels),
        reinterpret_cast<float4*>(out_pixels), batch_begin, count, num_samples, replace);
    WF_TRY(launch_check("reconstruct_shadow_kernel"));
    return t.finish();
}

extern "C" int wf_interpolate_colors(float* frame_pixels, float* pixels, int64_t num_pixels,
                                     int frame_index, float key_value, float white_point,
                                     wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(num_pixels >= 0, "wf_interpolate_colors: num_pixels < 0");
    if (num_pixels == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(frame_pixels && pixels, "wf_interpolate_colors: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    interpolate_kernel<<<grid_for(num_pixels, 256), 256, 0, s>>>(
        reinterpret_cast<float4*>(frame_pixels), reinterpret_cast<float4*>(pixels), num_pixels,
        frame_index, key_value, white_point);
    WF_TRY(launch_check("interpolate_kernel"));
    return t.finish();
}

extern "C" int wf_count_hits(const wf_hit* hits, int64_t n, uint32_t* d_count, wf_stream_t stream,
                             float* ms) {
    WF_CHECK_ARG(n >= 0 && d_count, "wf_count_hits: bad n or null counter");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(hits, "wf_count_hits: null buffer");
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    count_hits_kernel<<<(unsigned)std::min<int64_t>(grid_for(n, 256 * COUNT_HITS_ITEMS), COUNT_HITS_GRID),
                        256, 0, s>>>(hits, n, d_count);
    WF_TRY(launch_check("count_hits_kernel"));
    return t.finish();
}

// ------------------------------------------------------------ decision log
// The frame loop's discrete choices per path vertex, recorded by pixel for
// verification against an fp64 replay: the global triangle a ray hit (-1 for
// a miss) and, per shadow sample, whether the light was visible.
namespace wfpt {

__global__ __launch_bounds__(256) void log_hits_kernel(const wf_hit* __restrict__ hits,
                                                       const int32_t* __restrict__ s2i, int64_t n,
                                                       const int32_t* __restrict__ geom_first,
                                                       int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const wf_hit h = hits[i];
    out[s2i[i]] = h.prim < 0 ? -1 : geom_first[h.geom] + h.prim;
}

__global__ __launch_bounds__(256) void log_shadow_kernel(const wf_hit* __restrict__ shadow_hits,
                                                         const int32_t* __restrict__ shadow_i2s,
                                                         const int32_t* __restrict__ in_s2i,
                                                         int64_t begin, int64_t m, int S,
                                                         uint8_t* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m * S) return;
    const int64_t j = q / S;
    const int i = (int)(q - j * S);
    const int pixel = in_s2i[begin + j];
    out[(int64_t)pixel * S + i] = shadow_hits[shadow_i2s[q]].prim < 0 ? 1 : 0;
}

int log_hits(const wf_hit* hits, const int32_t* s2i, int64_t n, const int32_t* geom_first,
             int32_t* out, hipStream_t s) {
    if (n <= 0) return WF_OK;
    log_hits_kernel<<<grid_for(n, 256), 256, 0, s>>>(hits, s2i, n, geom_first, out);
    return launch_check("log_hits_kernel");
}

int log_shadow(const wf_hit* shadow_hits, const int32_t* shadow_i2s, const int32_t* in_s2i,
               int64_t begin, int64_t end, int S, uint8_t* out, hipStream_t s) {
    const int64_t m = end - begin;
    if (m <= 0) return WF_OK;
    log_shadow_kernel<<<grid_for(m * S, 256), 256, 0, s>>>(shadow_hits, shadow_i2s, in_s2i, begin,
                                                           m, S, out);
    return launch_check("log_shadow_kernel");
}

}  // namespace wfpt
