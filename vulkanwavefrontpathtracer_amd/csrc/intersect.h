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
};

__device__ __forceinline__ float pick3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Returns false for a zero direction (no triangle can be hit).
__device__ __forceinline__ bool wt_setup(float ox, float oy, float oz, float dx, float dy, float dz,
                                         WtRay& w) {
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

struct WtHit {
    float u, v, w, det, t;
};

// Edge functions and t; returns false on a miss.  t is not range checked.
__device__ __forceinline__ bool wt_intersect(const WtRay& r, float4 v0, float4 v1, float4 v2,
                                             WtHit& h) {
    const float a_x = pick3(v0.x, v0.y, v0.z, r.kx) - r.okx;
    const float a_y = pick3(v0.x, v0.y, v0.z, r.ky) - r.oky;
    const float a_z = pick3(v0.x, v0.y, v0.z, r.kz) - r.okz;
    const float b_x = pick3(v1.x, v1.y, v1.z, r.kx) - r.okx;
    const float b_y = pick3(v1.x, v1.y, v1.z, r.ky) - r.oky;
    const float b_z = pick3(v1.x, v1.y, v1.z, r.kz) - r.okz;
    const float c_x = pick3(v2.x, v2.y, v2.z, r.kx) - r.okx;
    const float c_y = pick3(v2.x, v2.y, v2.z, r.ky) - r.oky;
    const float c_z = pick3(v2.x, v2.y, v2.z, r.kz) - r.okz;
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
