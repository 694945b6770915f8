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

}  // namespace wfpt
