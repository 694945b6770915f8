// Device functions shared by the ray-generation, shade/extend and
// reconstruction kernels.  Each restates a GLSL helper of the reference; the
// CPU restatement mirrors the same operation order.
#pragma once

#include "scene.h"

namespace wfpt {

// initSeeds.comp:7-19 (TEA, 16 rounds in the reference's call).
__device__ __forceinline__ uint32_t tea(uint32_t v0, uint32_t v1, int rounds) {
    uint32_t s0 = 0u;
    for (int n = 0; n < rounds; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// raygenPath.comp:16-26 (LCG and its [0,1) float).
__device__ __forceinline__ float randf(uint32_t& seed) {
    seed = 1103515245u * seed + 12345u;
    return (float)(seed & 0x00FFFFFFu) / 16777216.0f;
}

// raygenShadow.comp:10-20 (Bob Jenkins' 96-bit mix).
__device__ __forceinline__ void jenkins_mix(uint32_t& a, uint32_t& b, uint32_t& c) {
    a -= b; a -= c; a ^= (c >> 13);
    b -= c; b -= a; b ^= (a << 8);
    c -= a; c -= b; c ^= (b >> 13);
    a -= b; a -= c; a ^= (c >> 12);
    b -= c; b -= a; b ^= (a << 16);
    c -= a; c -= b; c ^= (b >> 5);
    a -= b; a -= c; a ^= (c >> 3);
    b -= c; b -= a; b ^= (a << 10);
    c -= a; c -= b; c ^= (b >> 15);
}

// raygenShadow.comp:22-40: first two Sobol dimensions by XOR of matrix columns.
__device__ __forceinline__ void sobol2(uint32_t i, float& x, float& y) {
    uint32_t r1 = 0, r2 = 0;
    for (uint32_t v1 = 1u << 31, v2 = 3u << 30; i > 0u; i >>= 1) {
        if (i & 1u) {
            r1 ^= v1;
            r2 ^= v2 << 1;
        }
        v1 |= v1 >> 1;
        v2 ^= v2 >> 1;
    }
    x = (float)r1 / 4294967296.0f;
    y = (float)r2 / 4294967296.0f;
}

// raygenPrimary.comp:11-22 (radical inverse).
__device__ __forceinline__ float halton(int i, int base) {
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

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// normalize(v) = v * (1 / sqrt(dot(v, v)))
__device__ __forceinline__ V3 normalize3(V3 v) {
    float inv = 1.0f / sqrtf(dot3(v, v));
    return v3(v.x * inv, v.y * inv, v.z * inv);
}
// GLSL faceforward(N, I, Nref) with Nref = N.
__device__ __forceinline__ V3 faceforward3(V3 n, V3 i) {
    return dot3(n, i) < 0.0f ? n : v3(-n.x, -n.y, -n.z);
}

// Interpolated shading normal of triangle `tri` (geometrytypes.glsl:18-47:
// weights (1 - b.x - b.y, b.x, b.y) on the vertex normals, then normalize);
// returns the material index.
__device__ __forceinline__ int shading_normal(const float4* __restrict__ shade, int64_t tri,
                                              uint32_t packed_bary, V3& n) {
    float bu = (float)(packed_bary & 0xFFFFu) / 65535.0f;
    float bv = (float)(packed_bary >> 16) / 65535.0f;
    float bw = 1.0f - bu - bv;
    const float4 a = shade[3 * tri], b = shade[3 * tri + 1], c = shade[3 * tri + 2];
    V3 s = v3(a.x * bw + b.x * bu + c.x * bv, a.y * bw + b.y * bu + c.y * bv,
              a.z * bw + b.z * bu + c.z * bv);
    n = normalize3(s);
    return __float_as_int(a.w);
}

// ------------------------------------------------ per-ray stage functions
// Used by the wavefront kernels and by the path-replay adjoint kernel,
// which must reproduce every stage's arithmetic exactly.

struct Mat4 {
    float m[16];  // column-major (glm layout)
};

// raygenPrimary.comp:44-87: direction of the primary ray through pixel
// (x + hx, y + hy), screenToWorld * vec4(ndc, 0, 1) minus the camera.
__device__ __forceinline__ V3 primary_direction(const Mat4& s2w, int pixel, int width, int height,
                                                float hx, float hy, float ox, float oy, float oz) {
    float x = (float)(pixel % width) + hx;
    float y = (float)(pixel / width) + hy;
    float nx = 2.0f * x / (float)width - 1.0f;
    float ny = 2.0f * y / (float)height - 1.0f;
    const float* m = s2w.m;
    // mat4 * vec4(nx, ny, 0, 1), GLSL column-major
    float wx = ((m[0] * nx + m[4] * ny) + m[8] * 0.0f) + m[12] * 1.0f;
    float wy = ((m[1] * nx + m[5] * ny) + m[9] * 0.0f) + m[13] * 1.0f;
    float wz = ((m[2] * nx + m[6] * ny) + m[10] * 0.0f) + m[14] * 1.0f;
    float ww = ((m[3] * nx + m[7] * ny) + m[11] * 0.0f) + m[15] * 1.0f;
    V3 wp = v3(wx / ww, wy / ww, wz / ww);
    return normalize3(v3(wp.x - ox, wp.y - oy, wp.z - oz));
}

// raygenShadow.comp:91-97: per-path Cranley-Patterson offset from the
// Jenkins hash of (seed + key), key = the path's pixel.
__device__ __forceinline__ V3 shadow_offset(uint32_t seed, int key) {
    uint32_t ha = seed + (uint32_t)key, hb = 0x9e3779b9u, hc = 0x9e3779b9u;
    jenkins_mix(ha, hb, hc);
    jenkins_mix(ha, hb, hc);
    const float two_m32 = 2.3283064365386963e-10f;  // exp2(-32)
    return v3((float)ha * two_m32, (float)hb * two_m32, (float)hc * two_m32);
}

// raygenShadow.comp:110-133: shadow ray i of S from origin o towards the
// spherical-ish light volume; dead (tmax = -1) for a missed path ray.
__device__ __forceinline__ void shadow_ray(int i, int S, V3 off, V3 o, bool miss, float lx,
                                           float ly, float lz, float radius, float4& ra,
                                           float4& rb) {
    float px, py;
    sobol2((uint32_t)i, px, py);
    float pz = ((float)i + 0.5f) / (float)S;
    px += off.x;
    py += off.y;
    pz += off.z;
    if (px >= 1.0f) px -= 1.0f;
    if (py >= 1.0f) py -= 1.0f;
    if (pz >= 1.0f) pz -= 1.0f;
    px = 2.0f * px - 1.0f;
    py = 2.0f * py - 1.0f;
    pz = 2.0f * pz - 1.0f;
    V3 target = v3(lx + radius * px, ly + radius * py, lz + radius * pz);
    V3 dir = v3(target.x - o.x, target.y - o.y, target.z - o.z);
    float len = sqrtf(dot3(dir, dir));
    V3 nd = miss ? v3(1.f, 0.f, 0.f) : normalize3(dir);
    ra = make_float4(o.x, o.y, o.z, 1.0e-5f);
    rb = make_float4(nd.x, nd.y, nd.z, miss ? -1.0f : len);
}

// raygenShadow.comp:105-108: shadow origin backed off 1e-5 along the ray.
__device__ __forceinline__ V3 shadow_origin(float4 ro, float4 rd, float t, bool miss) {
    float adv = fmaxf(t - 1.0e-5f, 0.0f);
    V3 o = v3(ro.x + rd.x * adv, ro.y + rd.y * adv, ro.z + rd.z * adv);
    return miss ? v3(0.f, 0.f, 0.f) : o;
}

// raygenPath.comp:28-49: Duff et al. 2017 orthonormal basis and cosine lobe.
__device__ __forceinline__ V3 cosine_direction(float r0, float r1, V3 n) {
    const float two_pi = 2.0f * 3.14159265358979323846f;
    float theta = two_pi * r1;
    float radius = sqrtf(r0);
    float x = radius * sinf(theta);
    float z = radius * cosf(theta);
    float y = sqrtf(1.0f - r0);
    float s = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    V3 b1 = v3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    V3 b2 = v3(b, s + n.y * n.y * a, -n.y);
    return v3(x * b1.x + z * b2.x + y * n.x, x * b1.y + z * b2.y + y * n.y,
              x * b1.z + z * b2.z + y * n.z);
}

// reconstructSmooth.comp:76-84: Lambert term towards the
// point light, max(N.L, 0.1), with N faceforwarded against the ray.
__device__ __forceinline__ float smooth_alpha(V3 ff, V3 p, float lx, float ly, float lz) {
    V3 L = normalize3(v3(lx - p.x, ly - p.y, lz - p.z));
    return fmaxf(dot3(ff, L), 0.1f);
}

}  // namespace wfpt
