// Key-building device functions shared by the key kernels of the single-view
// stages and of the batched queue.  Conventions (mirrored exactly by the CPU
// restatement):
//   a    = clamp01((o - scene_min) / scene_size)   per component, IEEE div
//   norm = d * (1 / sqrt(dx*dx + dy*dy + dz*dz))    no fused multiply-add
//   clamp01(x) = fminf(fmaxf(x, 0), 1)              (NaN -> 0)
//   quantisation truncates: uint(x * scale)
#pragma once

#include "common.h"

namespace wfpt {

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// Spread the low 10 bits of v so that bit k lands on bit 3k.
__device__ __forceinline__ uint32_t spread3_10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
// Spread the low 16 bits of v so that bit k lands on bit 2k.
__device__ __forceinline__ uint32_t spread2_16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

struct KeyFrame {
    float minx, miny, minz, sx, sy, sz;
};

__device__ __forceinline__ void origin_unit(const KeyFrame& f, float4 o, float& ax, float& ay,
                                            float& az) {
    ax = clamp01((o.x - f.minx) / f.sx);
    ay = clamp01((o.y - f.miny) / f.sy);
    az = clamp01((o.z - f.minz) / f.sz);
}

// 10-bit origin coordinates within the scene box.
__device__ __forceinline__ void origin_q10(const KeyFrame& f, float4 o, uint32_t& ix, uint32_t& iy,
                                           uint32_t& iz) {
    float ax, ay, az;
    origin_unit(f, o, ax, ay, az);
    ix = (uint32_t)(ax * 1023.0f);
    iy = (uint32_t)(ay * 1023.0f);
    iz = (uint32_t)(az * 1023.0f);
}

__device__ __forceinline__ uint32_t origin30(const KeyFrame& f, float4 o) {
    uint32_t ix, iy, iz;
    origin_q10(f, o, ix, iy, iz);
    return (spread3_10(ix) << 2) | (spread3_10(iy) << 1) | spread3_10(iz);
}

// Octahedral direction coordinates (Cigolle et al. 2014), `bits` bits each.
__device__ __forceinline__ void oct_q(float4 d, int bits, uint32_t& qu, uint32_t& qv) {
    float l1 = (fabsf(d.x) + fabsf(d.y)) + fabsf(d.z);
    float px = d.x / l1, py = d.y / l1;
    if (d.z < 0.0f) {
        float nx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = nx;
        py = ny;
    }
    float u = px * 0.5f + 0.5f, v = py * 0.5f + 0.5f;
    float scale = (float)((1u << bits) - 1u);
    qu = (uint32_t)(clamp01(u) * scale);
    qv = (uint32_t)(clamp01(v) * scale);
}

// Octahedral direction code with `bits` bits per axis, u-major 2-D Morton.
__device__ __forceinline__ uint32_t oct_code(float4 d, int bits) {
    uint32_t qu, qv;
    oct_q(d, bits, qu, qv);
    return (spread2_16(qu) << 1) | spread2_16(qv);
}

// Spread the low 10 bits of v so that bit k lands on bit 5k.  Bit k must move
// left by 4k = 32a + 16b + 8c + 4d for k = 8a + 4b + 2c + d, so four
// shift-and-mask steps (by 32, 16, 8, 4) move every bit whose k has that
// binary digit set.
__device__ __forceinline__ uint64_t spread5_10(uint32_t v) {
    uint64_t x = v & 0x3ffu;
    x = (x | (x << 32)) & 0x00000300000000ffull;
    x = (x | (x << 16)) & 0x0000030000f0000full;
    x = (x | (x << 8)) & 0x00000300c0300c03ull;
    x = (x | (x << 4)) & 0x0000210842108421ull;
    return x;
}

// 5-D Morton code of the same 30 origin and 30 octahedral-direction bits as
// the 60-bit key: 10 levels of (u, v, x, y, z) bits from the top (direction
// bits 14..5 beside origin bits 9..0), then the direction's low 5 levels as
// (u, v) pairs.  Rays adjacent in this order share a coarse origin cell and
// a coarse direction cone at every scale, which is what keeps a wave's
// traversals together; the concatenated key groups by origin only.
// DBITS = 15 gives the 60-bit key, 14 the 58 bits below the batched key's
// 6-bit scene prefix (4 low direction levels).
template <int DBITS>
__device__ __forceinline__ uint64_t morton5d(const KeyFrame& f, float4 o, float4 d) {
    constexpr int LOW = DBITS - 10;  // direction levels below the origin's
    uint32_t ix, iy, iz, qu, qv;
    origin_q10(f, o, ix, iy, iz);
    oct_q(d, DBITS, qu, qv);
    const uint64_t hi = (spread5_10(qu >> LOW) << 4) | (spread5_10(qv >> LOW) << 3) |
                        (spread5_10(ix) << 2) | (spread5_10(iy) << 1) | spread5_10(iz);
    const uint32_t m = (1u << LOW) - 1u;
    const uint32_t lo = (spread2_16(qu & m) << 1) | spread2_16(qv & m);
    return (hi << (2 * LOW)) | lo;
}
__device__ __forceinline__ uint64_t morton5d60(const KeyFrame& f, float4 o, float4 d) {
    return morton5d<15>(f, o, d);
}
// Its top 30 bits as a u32 key: 6 levels of (u, v, x, y, z), direction bits
// 14..9 beside origin bits 9..4.  Sorted in 4 passes of 8-bit digits instead
// of 8; rays equal in these bits keep their order (the sort is stable).
__device__ __forceinline__ uint32_t morton5d30(const KeyFrame& f, float4 o, float4 d) {
    return (uint32_t)(morton5d<15>(f, o, d) >> 30);
}

__device__ __forceinline__ uint32_t aila32(const KeyFrame& f, float4 o, float4 d) {
    float ax, ay, az;
    origin_unit(f, o, ax, ay, az);
    float inv = 1.0f / sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    float bx = clamp01((d.x * inv + 1.0f) * 0.5f);
    float by = clamp01((d.y * inv + 1.0f) * 0.5f);
    float bz = clamp01((d.z * inv + 1.0f) * 0.5f);
    uint32_t ia[3] = {(uint32_t)(ax * 127.0f), (uint32_t)(ay * 127.0f), (uint32_t)(az * 127.0f)};
    uint32_t ib[3] = {(uint32_t)(bx * 15.0f), (uint32_t)(by * 15.0f), (uint32_t)(bz * 15.0f)};
    // Triples of (x,y,z) bits from the most significant end: origin bits
    // 6,5,4,3, direction bit 3, origin 2, direction 2, origin 1, direction 1,
    // origin 0, then the x and y direction bits 0.
    auto tri = [](const uint32_t* v, int k) {
        return (((v[0] >> k) & 1u) << 2) | (((v[1] >> k) & 1u) << 1) | ((v[2] >> k) & 1u);
    };
    uint32_t key = 0;
    key |= tri(ia, 6) << 29;
    key |= tri(ia, 5) << 26;
    key |= tri(ia, 4) << 23;
    key |= tri(ia, 3) << 20;
    key |= tri(ib, 3) << 17;
    key |= tri(ia, 2) << 14;
    key |= tri(ib, 2) << 11;
    key |= tri(ia, 1) << 8;
    key |= tri(ib, 1) << 5;
    key |= tri(ia, 0) << 2;
    key |= ((ib[0] & 1u) << 1) | (ib[1] & 1u);
    return key;
}

}  // namespace wfpt
