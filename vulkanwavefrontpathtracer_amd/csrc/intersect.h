// Watertight ray/triangle intersection (Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2013) in fp32 with the paper's fp64
// fallback when an edge function is exactly zero.  The CPU restatement
// performs the same operations in the same order, so hit decisions, t and
// barycentrics agree bit for bit.
#pragma once

#include "common.h"

namespace wfpt {

struct WtRay {
    int kx, ky, kz;
    float sx, sy, sz;  // shear constants
    float okx, oky, okz;
    // Axis permutation kz * 2 + (d[kz] < 0) when every ray of the wave has
    // the same one (wave-uniform; set by wt_uniform_perm), else -1.  The
    // intersection then takes the coordinates in a fixed order instead of
    // selecting them per lane; the arithmetic is the same.
    int uperm;
};

__device__ __forceinline__ float pick3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Returns false for a zero direction (no triangle can be hit).
__device__ __forceinline__ bool wt_setup(float ox, float oy, float oz, float dx, float dy, float dz,
                                         WtRay& w) {
    w.uperm = -1;
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = pick3(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    if (!(dkz != 0.0f)) return false;
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.sx = pick3(dx, dy, dz, kx) / dkz;
    w.sy = pick3(dx, dy, dz, ky) / dkz;
    w.sz = 1.0f / dkz;
    w.okx = pick3(ox, oy, oz, kx);
    w.oky = pick3(ox, oy, oz, ky);
    w.okz = pick3(ox, oy, oz, kz);
    return true;
}

// Marks the rays of a wave whose live lanes share one axis permutation
// (called by every lane of the wave; dead lanes do not count).
__device__ __forceinline__ void wt_uniform_perm(WtRay& w, bool live) {
    const int perm = w.kz * 2 + (w.kx == (w.kz == 2 ? 0 : w.kz + 1) ? 0 : 1);
    const uint64_t lm = ballot64(live);
    if (lm == 0) {
        w.uperm = -1;
        return;
    }
    const int first = __ffsll((unsigned long long)lm) - 1;
    const int p0 = __builtin_amdgcn_readlane(perm, first);
    w.uperm = __all(!live || perm == p0) ? p0 : -1;
}

struct WtHit {
    float u, v, w, det, t;
};

__device__ __forceinline__ bool wt_edges(const WtRay& r, float a_x, float a_y, float a_z, float b_x,
                                         float b_y, float b_z, float c_x, float c_y, float c_z,
                                         WtHit& h);

// Coordinates of v in the order (kx, ky, kz) of permutation P (compile time).
template <int P>
__device__ __forceinline__ void perm3(float4 v, float& x, float& y, float& z) {
    constexpr int KZ = P / 2;
    constexpr int KX0 = KZ == 2 ? 0 : KZ + 1;
    constexpr int KY0 = KX0 == 2 ? 0 : KX0 + 1;
    constexpr int KX = (P & 1) ? KY0 : KX0, KY = (P & 1) ? KX0 : KY0;
    const float c[3] = {v.x, v.y, v.z};
    x = c[KX];
    y = c[KY];
    z = c[KZ];
}

template <int P>
__device__ __forceinline__ bool wt_intersect_p(const WtRay& r, float4 v0, float4 v1, float4 v2,
                                               WtHit& h) {
    float x0, y0, z0, x1, y1, z1, x2, y2, z2;
    perm3<P>(v0, x0, y0, z0);
    perm3<P>(v1, x1, y1, z1);
    perm3<P>(v2, x2, y2, z2);
    return wt_edges(r, x0 - r.okx, y0 - r.oky, z0 - r.okz, x1 - r.okx, y1 - r.oky, z1 - r.okz,
                    x2 - r.okx, y2 - r.oky, z2 - r.okz, h);
}

// Edge functions and t; returns false on a miss.  t is not range checked.
__device__ __forceinline__ bool wt_intersect(const WtRay& r, float4 v0, float4 v1, float4 v2,
                                             WtHit& h) {
    switch (r.uperm) {
        case 0: return wt_intersect_p<0>(r, v0, v1, v2, h);
        case 1: return wt_intersect_p<1>(r, v0, v1, v2, h);
        case 2: return wt_intersect_p<2>(r, v0, v1, v2, h);
        case 3: return wt_intersect_p<3>(r, v0, v1, v2, h);
        case 4: return wt_intersect_p<4>(r, v0, v1, v2, h);
        case 5: return wt_intersect_p<5>(r, v0, v1, v2, h);
        default: break;
    }
    const float a_x = pick3(v0.x, v0.y, v0.z, r.kx) - r.okx;
    const float a_y = pick3(v0.x, v0.y, v0.z, r.ky) - r.oky;
    const float a_z = pick3(v0.x, v0.y, v0.z, r.kz) - r.okz;
    const float b_x = pick3(v1.x, v1.y, v1.z, r.kx) - r.okx;
    const float b_y = pick3(v1.x, v1.y, v1.z, r.ky) - r.oky;
    const float b_z = pick3(v1.x, v1.y, v1.z, r.kz) - r.okz;
    const float c_x = pick3(v2.x, v2.y, v2.z, r.kx) - r.okx;
    const float c_y = pick3(v2.x, v2.y, v2.z, r.ky) - r.oky;
    const float c_z = pick3(v2.x, v2.y, v2.z, r.kz) - r.okz;
    return wt_edges(r, a_x, a_y, a_z, b_x, b_y, b_z, c_x, c_y, c_z, h);
}

// The watertight test on the ray-space coordinates (a, b, c) of the three
// vertices relative to the origin, in the order (kx, ky, kz).
__device__ __forceinline__ bool wt_edges(const WtRay& r, float a_x, float a_y, float a_z, float b_x,
                                         float b_y, float b_z, float c_x, float c_y, float c_z,
                                         WtHit& h) {
    const float Ax = fmaf(-r.sx, a_z, a_x), Ay = fmaf(-r.sy, a_z, a_y);
    const float Bx = fmaf(-r.sx, b_z, b_x), By = fmaf(-r.sy, b_z, b_y);
    const float Cx = fmaf(-r.sx, c_z, c_x), Cy = fmaf(-r.sy, c_z, c_y);
    float U = fmaf(Cx, By, -(Cy * Bx));
    float V = fmaf(Ax, Cy, -(Ay * Cx));
    float W = fmaf(Bx, Ay, -(By * Ax));
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    const float Az = r.sz * a_z, Bz = r.sz * b_z, Cz = r.sz * c_z;
    const float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    h.u = U;
    h.v = V;
    h.w = W;
    h.det = det;
    h.t = T / det;
    return true;
}

// packUnorm2x16 with round-half-to-even: low 16 bits = first component.
__device__ __forceinline__ uint32_t pack_unorm2x16(float a, float b) {
    float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
    uint32_t ua = (uint32_t)rintf(ca * 65535.0f), ub = (uint32_t)rintf(cb * 65535.0f);
    return ua | (ub << 16);
}

}  // namespace wfpt
