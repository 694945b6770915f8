// Ray sort keys.
//
// WF_KEY_AILA32 restates computeMortonCodesAila32.comp:39-88 (7 origin bits
// and 4 direction bits per axis, the direction's z bit 0 dropped).  The other
// modes are this library's own (origin-only 30-bit key, and 60-/64-bit keys
// with an octahedral direction code after Cigolle et al. 2014, concatenated
// after the origin code or interleaved with it as a 5-D Morton code).
//
// Arithmetic conventions (mirrored exactly by the CPU restatement):
//   a    = clamp01((o - scene_min) / scene_size)   per component, IEEE div
//   norm = d * (1 / sqrt(dx*dx + dy*dy + dz*dz))    no fused multiply-add
//   clamp01(x) = fminf(fmaxf(x, 0), 1)              (NaN -> 0)
//   quantisation truncates: uint(x * scale)
#include "keys.h"

namespace wfpt {

template <int MODE>
__global__ __launch_bounds__(256) void keys_kernel(const wf_ray* __restrict__ rays,
                                                   void* __restrict__ keys,
                                                   uint32_t* __restrict__ indices, int64_t n,
                                                   KeyFrame f) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 o, d;
    load_ray(rays, i, o, d);
    if (MODE == WF_KEY_AILA32) {
        static_cast<uint32_t*>(keys)[i] = aila32(f, o, d);
    } else if (MODE == WF_KEY_ORIGIN30) {
        static_cast<uint32_t*>(keys)[i] = origin30(f, o);
    } else if (MODE == WF_KEY_ORIGIN_DIR_MORTON60) {
        static_cast<uint64_t*>(keys)[i] = morton5d60(f, o, d);
    } else if (MODE == WF_KEY_ORIGIN_DIR_MORTON30) {
        static_cast<uint32_t*>(keys)[i] = morton5d30(f, o, d);
    } else {
        uint64_t k = ((uint64_t)origin30(f, o) << 30) | (uint64_t)oct_code(d, 15);
        static_cast<uint64_t*>(keys)[i] = k;
    }
    indices[i] = (uint32_t)i;
}

struct SceneFrames {
    float f[6 * 64];  // per scene: min.xyz, size.xyz
};

__global__ __launch_bounds__(256) void keys_batched_kernel(
    const wf_ray* __restrict__ rays, const uint8_t* __restrict__ scene_of_ray,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ indices, int64_t n,
    const SceneFrames frames) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 o, d;
    load_ray(rays, i, o, d);
    uint32_t s = scene_of_ray[i];
    const float* fr = frames.f + 6 * (s & 63u);
    KeyFrame f{fr[0], fr[1], fr[2], fr[3], fr[4], fr[5]};
    uint64_t k = ((uint64_t)(s & 63u) << 58) | ((uint64_t)origin30(f, o) << 28) |
                 (uint64_t)oct_code(d, 14);
    keys[i] = k;
    indices[i] = (uint32_t)i;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_key_bytes(int mode) {
    switch (mode) {
        case WF_KEY_AILA32:
        case WF_KEY_ORIGIN30:
        case WF_KEY_ORIGIN_DIR_MORTON30: return 4;
        case WF_KEY_ORIGIN30_OCT30:
        case WF_KEY_SCENE6_ORIGIN30_OCT28:
        case WF_KEY_ORIGIN_DIR_MORTON60: return 8;
        default: return WF_ERR_INVALID_ARG;
    }
}

extern "C" int wf_key_bits(int mode) {
    switch (mode) {
        case WF_KEY_AILA32: return 32;
        case WF_KEY_ORIGIN30:
        case WF_KEY_ORIGIN_DIR_MORTON30: return 30;
        case WF_KEY_ORIGIN30_OCT30:
        case WF_KEY_ORIGIN_DIR_MORTON60: return 60;
        case WF_KEY_SCENE6_ORIGIN30_OCT28: return 64;
        default: return WF_ERR_INVALID_ARG;
    }
}

extern "C" int wf_morton_keys(const wf_ray* rays, void* keys, uint32_t* indices, int64_t n,
                              int mode, const float scene_min[3], const float scene_max[3],
                              wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(n >= 0, "wf_morton_keys: n < 0");
    WF_CHECK_ARG((mode >= WF_KEY_AILA32 && mode <= WF_KEY_ORIGIN30_OCT30) ||
                     mode == WF_KEY_ORIGIN_DIR_MORTON60 || mode == WF_KEY_ORIGIN_DIR_MORTON30,
                 "wf_morton_keys: unknown key mode (use wf_morton_keys_batched for scene keys)");
    WF_CHECK_ARG(scene_min && scene_max, "wf_morton_keys: null scene bounds");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(rays && keys && indices, "wf_morton_keys: null buffer");
    WF_CHECK_ARG(n <= (int64_t)UINT32_MAX, "wf_morton_keys: n exceeds u32 indices");
    // Tracer::computeMortonCodes passes sceneSize = max - min (Tracer.cpp:29).
    KeyFrame f{scene_min[0], scene_min[1], scene_min[2], scene_max[0] - scene_min[0],
               scene_max[1] - scene_min[1], scene_max[2] - scene_min[2]};
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    unsigned g = grid_for(n, 256);
    switch (mode) {
        case WF_KEY_AILA32:
            keys_kernel<WF_KEY_AILA32><<<g, 256, 0, s>>>(rays, keys, indices, n, f);
            break;
        case WF_KEY_ORIGIN30:
            keys_kernel<WF_KEY_ORIGIN30><<<g, 256, 0, s>>>(rays, keys, indices, n, f);
            break;
        case WF_KEY_ORIGIN_DIR_MORTON60:
            keys_kernel<WF_KEY_ORIGIN_DIR_MORTON60><<<g, 256, 0, s>>>(rays, keys, indices, n, f);
            break;
        case WF_KEY_ORIGIN_DIR_MORTON30:
            keys_kernel<WF_KEY_ORIGIN_DIR_MORTON30><<<g, 256, 0, s>>>(rays, keys, indices, n, f);
            break;
        default:
            keys_kernel<WF_KEY_ORIGIN30_OCT30><<<g, 256, 0, s>>>(rays, keys, indices, n, f);
            break;
    }
    WF_TRY(launch_check("keys_kernel"));
    return t.finish();
}

extern "C" int wf_morton_keys_batched(const wf_ray* rays, const uint8_t* scene_of_ray,
                                      void* keys, uint32_t* indices, int64_t n,
                                      const float* scene_min, const float* scene_max,
                                      int num_scenes, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(n >= 0, "wf_morton_keys_batched: n < 0");
    WF_CHECK_ARG(num_scenes >= 1 && num_scenes <= 64,
                 "wf_morton_keys_batched: num_scenes must be in [1, 64]");
    WF_CHECK_ARG(scene_min && scene_max, "wf_morton_keys_batched: null scene bounds");
    if (n == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(rays && keys && indices && scene_of_ray, "wf_morton_keys_batched: null buffer");
    WF_CHECK_ARG(n <= (int64_t)UINT32_MAX, "wf_morton_keys_batched: n exceeds u32 indices");
    SceneFrames frames{};
    for (int k = 0; k < num_scenes; ++k)
        for (int a = 0; a < 3; ++a) {
            frames.f[6 * k + a] = scene_min[3 * k + a];
            frames.f[6 * k + 3 + a] = scene_max[3 * k + a] - scene_min[3 * k + a];
        }
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    keys_batched_kernel<<<grid_for(n, 256), 256, 0, s>>>(
        rays, scene_of_ray, static_cast<uint64_t*>(keys), indices, n, frames);
    WF_TRY(launch_check("keys_batched_kernel"));
    return t.finish();
}
