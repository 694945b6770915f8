/*
 * wfpt_oracle.c - CPU restatement of the reference wavefront path tracer's
 * hot path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark call it as the checker; no product code
 * links or calls it.
 *
 * Compiled with gcc -O2 -ffp-contract=off -fno-fast-math so that every fp32
 * expression is evaluated in the written order with IEEE rounding, matching
 * the device code, which is built with the same contraction rule.
 *
 * Parity pins (tests/test_oracle.py):
 *   keys    - hand-computed known-answer vectors + brute-force bit checks.
 *   sort    - stable merge sort (an algorithm independent of the GPU's LSD
 *             radix sort) checked against numpy's stable argsort.
 *   reorder - permutation invariants.
 *   trace   - FP64 brute force; FP32 watertight test checked against the
 *             FP64 one on random and constructed edge cases.
 * Functions without such a check say "parity unpinned" in their comment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

static float clamp01f(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

/* ----------------------------------------------------------------- keys */

/* computeMortonCodesAila32.comp:39-40, :42-43 (quantisation). */
static void aila_quantise(const float* ray, const float mn[3], const float sz[3],
                          uint32_t ia[3], uint32_t ib[3]) {
    float a[3], b[3];
    for (int k = 0; k < 3; ++k) a[k] = clamp01f((ray[k] - mn[k]) / sz[k]);
    /* normalize(d) = d * (1 / sqrt(dot(d, d))) */
    float d2 = ray[4] * ray[4] + ray[5] * ray[5] + ray[6] * ray[6];
    float inv = 1.0f / sqrtf(d2);
    for (int k = 0; k < 3; ++k) b[k] = clamp01f((ray[4 + k] * inv + 1.0f) * 0.5f);
    for (int k = 0; k < 3; ++k) {
        ia[k] = (uint32_t)(a[k] * 127.0f);
        ib[k] = (uint32_t)(b[k] * 15.0f);
    }
}

/* The bit assignment of computeMortonCodesAila32.comp:47-88 as a table:
 * {source (0 = origin ia, 1 = direction ib), axis, source bit, key bit}. */
static const int AILA_TABLE[32][4] = {
    {0, 0, 6, 31}, {0, 1, 6, 30}, {0, 2, 6, 29}, {0, 0, 5, 28}, {0, 1, 5, 27}, {0, 2, 5, 26},
    {0, 0, 4, 25}, {0, 1, 4, 24}, {0, 2, 4, 23}, {0, 0, 3, 22}, {0, 1, 3, 21}, {0, 2, 3, 20},
    {1, 0, 3, 19}, {1, 1, 3, 18}, {1, 2, 3, 17}, {0, 0, 2, 16}, {0, 1, 2, 15}, {0, 2, 2, 14},
    {1, 0, 2, 13}, {1, 1, 2, 12}, {1, 2, 2, 11}, {0, 0, 1, 10}, {0, 1, 1, 9},  {0, 2, 1, 8},
    {1, 0, 1, 7},  {1, 1, 1, 6},  {1, 2, 1, 5},  {0, 0, 0, 4},  {0, 1, 0, 3},  {0, 2, 0, 2},
    {1, 0, 0, 1},  {1, 1, 0, 0}};

OR_API void or_keys_aila32(const float* rays, int64_t n, const float mn[3], const float mx[3],
                           uint32_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    for (int64_t i = 0; i < n; ++i) {
        uint32_t ia[3], ib[3];
        aila_quantise(rays + 8 * i, mn, sz, ia, ib);
        uint32_t key = 0;
        for (int r = 0; r < 32; ++r) {
            const int* e = AILA_TABLE[r];
            uint32_t v = e[0] == 0 ? ia[e[1]] : ib[e[1]];
            key |= ((v >> e[2]) & 1u) << e[3];
        }
        keys[i] = key;
    }
}

/* 10-bit-per-axis origin Morton code: bit 3k+2 = x bit k, 3k+1 = y, 3k = z. */
static uint32_t origin30(const float* ray, const float mn[3], const float sz[3]) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = (uint32_t)(clamp01f((ray[k] - mn[k]) / sz[k]) * 1023.0f);
    uint32_t key = 0;
    for (int b = 0; b < 10; ++b)
        for (int k = 0; k < 3; ++k) key |= ((q[k] >> b) & 1u) << (3 * b + (2 - k));
    return key;
}

/* Octahedral map (Cigolle et al. 2014) of the direction, `bits` bits per
 * axis, 2-D Morton with u above v. */
static uint32_t oct_code(const float* ray, int bits) {
    float dx = ray[4], dy = ray[5], dz = ray[6];
    float l1 = (fabsf(dx) + fabsf(dy)) + fabsf(dz);
    float px = dx / l1, py = dy / l1;
    if (dz < 0.0f) {
        float nx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float ny = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = nx;
        py = ny;
    }
    float u = px * 0.5f + 0.5f, v = py * 0.5f + 0.5f;
    float scale = (float)((1u << bits) - 1u);
    uint32_t qu = (uint32_t)(clamp01f(u) * scale), qv = (uint32_t)(clamp01f(v) * scale);
    uint32_t code = 0;
    for (int b = 0; b < bits; ++b) code |= (((qu >> b) & 1u) << (2 * b + 1)) | (((qv >> b) & 1u) << (2 * b));
    return code;
}

OR_API void or_keys_origin30(const float* rays, int64_t n, const float mn[3], const float mx[3],
                             uint32_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    for (int64_t i = 0; i < n; ++i) keys[i] = origin30(rays + 8 * i, mn, sz);
}

OR_API void or_keys_oct60(const float* rays, int64_t n, const float mn[3], const float mx[3],
                          uint64_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    for (int64_t i = 0; i < n; ++i)
        keys[i] = ((uint64_t)origin30(rays + 8 * i, mn, sz) << 30) | oct_code(rays + 8 * i, 15);
}

OR_API void or_keys_scene64(const float* rays, const uint8_t* scene, int64_t n, const float* mn,
                            const float* mx, uint64_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        int s = scene[i] & 63;
        float sz[3] = {mx[3 * s] - mn[3 * s], mx[3 * s + 1] - mn[3 * s + 1], mx[3 * s + 2] - mn[3 * s + 2]};
        keys[i] = ((uint64_t)s << 58) | ((uint64_t)origin30(rays + 8 * i, mn + 3 * s, sz) << 28) |
                  oct_code(rays + 8 * i, 14);
    }
}

/* ----------------------------------------------------------------- sort */
/* Stable ascending sort of (key & mask) with values, by bottom-up merge sort
 * (vk_radix_sort.h:2598-2755 semantics: stable key-value sort, in place). */
OR_API void or_sort_pairs(void* keys, uint32_t* vals, int64_t n, int key_bytes, int key_bits) {
    if (n <= 1) return;
    uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
    uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* v2 = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint64_t* full = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint64_t* full2 = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* v = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (int64_t i = 0; i < n; ++i) {
        full[i] = key_bytes == 4 ? ((uint32_t*)keys)[i] : ((uint64_t*)keys)[i];
        k[i] = full[i] & mask;
        v[i] = vals[i];
    }
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                if (k[b] < k[a]) { k2[o] = k[b]; v2[o] = v[b]; full2[o++] = full[b++]; }
                else { k2[o] = k[a]; v2[o] = v[a]; full2[o++] = full[a++]; }
            }
            while (a < mid) { k2[o] = k[a]; v2[o] = v[a]; full2[o++] = full[a++]; }
            while (b < hi) { k2[o] = k[b]; v2[o] = v[b]; full2[o++] = full[b++]; }
        }
        uint64_t* t = k; k = k2; k2 = t;
        uint32_t* tv = v; v = v2; v2 = tv;
        uint64_t* tf = full; full = full2; full2 = tf;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (key_bytes == 4) ((uint32_t*)keys)[i] = (uint32_t)full[i];
        else ((uint64_t*)keys)[i] = full[i];
        vals[i] = v[i];
    }
    free(k); free(k2); free(v2); free(full); free(full2); free(v);
}

/* -------------------------------------------------------------- reorder */
/* reorderRays.comp:37-44. */
OR_API void or_reorder(const uint32_t* idx, const float* in_rays, float* out_rays,
                       const int32_t* in_s2i, int32_t* out_s2i, int32_t* out_i2s, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t src = idx[i];
        memcpy(out_rays + 8 * i, in_rays + 8 * (int64_t)src, 32);
        int32_t index = in_s2i[src];
        out_s2i[i] = index;
        out_i2s[index] = (int32_t)i;
    }
}

/* ---------------------------------------------------------------- trace */
/* trace.comp:27-88 semantics with the driver's traversal replaced by a
 * brute-force scan: opaque, no culling, interval [tmin, tmax], closest hit
 * with ties on t broken by the smaller triangle index, or any hit.
 * The fp32 test restates Woop, Benthin & Wald 2013 (watertight) with the
 * same operation order as the device code; the fp64 test below is an
 * independent textbook intersector used to pin it. */

typedef struct { int kx, ky, kz; float sx, sy, sz, okx, oky, okz; } wt_ray;

static float pick3(const float* v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

static int wt_setup(const float* o, const float* d, wt_ray* w) {
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = pick3(d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    if (!(dkz != 0.0f)) return 0;
    w->kx = kx; w->ky = ky; w->kz = kz;
    w->sx = pick3(d, kx) / dkz;
    w->sy = pick3(d, ky) / dkz;
    w->sz = 1.0f / dkz;
    w->okx = pick3(o, kx); w->oky = pick3(o, ky); w->okz = pick3(o, kz);
    return 1;
}

/* returns 1 on a hit of the infinite-interval test; writes t, U, V, W, det */
static int wt_intersect(const wt_ray* r, const float* v0, const float* v1, const float* v2,
                        float* t, float* U_, float* V_, float* W_, float* det_) {
    float a_x = pick3(v0, r->kx) - r->okx, a_y = pick3(v0, r->ky) - r->oky, a_z = pick3(v0, r->kz) - r->okz;
    float b_x = pick3(v1, r->kx) - r->okx, b_y = pick3(v1, r->ky) - r->oky, b_z = pick3(v1, r->kz) - r->okz;
    float c_x = pick3(v2, r->kx) - r->okx, c_y = pick3(v2, r->ky) - r->oky, c_z = pick3(v2, r->kz) - r->okz;
    float Ax = fmaf(-r->sx, a_z, a_x), Ay = fmaf(-r->sy, a_z, a_y);
    float Bx = fmaf(-r->sx, b_z, b_x), By = fmaf(-r->sy, b_z, b_y);
    float Cx = fmaf(-r->sx, c_z, c_x), Cy = fmaf(-r->sy, c_z, c_y);
    float U = fmaf(Cx, By, -(Cy * Bx));
    float V = fmaf(Ax, Cy, -(Ay * Cx));
    float W = fmaf(Bx, Ay, -(By * Ax));
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    float det = (U + V) + W;
    if (det == 0.0f) return 0;
    float Az = r->sz * a_z, Bz = r->sz * b_z, Cz = r->sz * c_z;
    float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    *t = T / det; *U_ = U; *V_ = V; *W_ = W; *det_ = det;
    return 1;
}

static uint32_t pack_unorm2x16(float a, float b) {
    float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
    return (uint32_t)rintf(ca * 65535.0f) | ((uint32_t)rintf(cb * 65535.0f) << 16);
}

/* tri_verts: 9 floats per triangle (original order).  out: per ray
 * {triangle index or -1, t bits, packed bary, 0} as int32[4]. */
OR_API void or_trace_f32(const float* rays, int64_t n, const float* tri_verts, int64_t ntri,
                         int closest, int32_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float* ry = rays + 8 * i;
        float tmin = ry[3], tmax = ry[7];
        int32_t best = -1;
        float best_t = tmax, bu = 0, bv = 0, bw = 0, bd = 1;
        wt_ray w;
        if (tmax >= tmin && wt_setup(ry, ry + 4, &w)) {
            for (int64_t k = 0; k < ntri; ++k) {
                float t, U, V, W, det;
                const float* tv = tri_verts + 9 * k;
                if (!wt_intersect(&w, tv, tv + 3, tv + 6, &t, &U, &V, &W, &det)) continue;
                if (!(t >= tmin && t <= best_t)) continue;
                if (closest) {
                    if (t < best_t || best < 0 || k < best) {
                        if (t == best_t && best >= 0 && k > best) continue;
                        best = (int32_t)k; best_t = t; bu = U; bv = V; bw = W; bd = det;
                    }
                } else {
                    best = (int32_t)k; best_t = t; bu = U; bv = V; bw = W; bd = det;
                    break;
                }
            }
        }
        int32_t* o = out + 4 * i;
        if (best >= 0) {
            float tb = best_t;
            o[0] = best;
            memcpy(&o[1], &tb, 4);
            o[2] = (int32_t)pack_unorm2x16(bv / bd, bw / bd);
            (void)bu;
        } else {
            float m1 = -1.0f;
            o[0] = -1;
            memcpy(&o[1], &m1, 4);
            o[2] = 0;
        }
        o[3] = 0;
    }
}

/* Independent fp64 intersector (Moller-Trumbore in double, two-sided).
 * out_t[i] = closest t in [tmin, tmax] or -1, out_tri[i] = triangle index
 * (smallest index among exact-t ties), out_edge[i] = smallest barycentric
 * coordinate of that hit (distance to the nearest edge in barycentric
 * units), used to recognise grazing-edge cases. */
OR_API void or_trace_f64(const float* rays, int64_t n, const float* tri_verts, int64_t ntri,
                         double* out_t, int32_t* out_tri, double* out_edge) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float* ry = rays + 8 * i;
        double o[3] = {ry[0], ry[1], ry[2]}, d[3] = {ry[4], ry[5], ry[6]};
        double tmin = ry[3], tmax = ry[7];
        double best_t = tmax;
        int32_t best = -1;
        double best_edge = 0;
        if (tmax >= tmin) {
            for (int64_t k = 0; k < ntri; ++k) {
                const float* tv = tri_verts + 9 * k;
                double v0[3] = {tv[0], tv[1], tv[2]};
                double e1[3] = {tv[3] - v0[0], tv[4] - v0[1], tv[5] - v0[2]};
                double e2[3] = {tv[6] - v0[0], tv[7] - v0[1], tv[8] - v0[2]};
                double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
                double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
                if (det == 0.0) continue;
                double inv = 1.0 / det;
                double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
                double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
                if (u < 0.0 || u > 1.0) continue;
                double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
                double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
                if (v < 0.0 || u + v > 1.0) continue;
                double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
                if (!(t >= tmin && t <= best_t)) continue;
                if (t == best_t && best >= 0 && k > best) continue;
                best = (int32_t)k; best_t = t;
                double w = 1.0 - u - v;
                best_edge = fmin(fmin(u, v), w);
            }
        }
        out_t[i] = best >= 0 ? best_t : -1.0;
        out_tri[i] = best;
        out_edge[i] = best_edge;
    }
}

/* Smallest barycentric coordinate (fp64) of ray i against triangle k, or
 * -1 if the ray's line is parallel to it; used to classify disagreements. */
OR_API double or_edge_distance_f64(const float* ray, const float* tv) {
    double o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
    double v0[3] = {tv[0], tv[1], tv[2]};
    double e1[3] = {tv[3] - v0[0], tv[4] - v0[1], tv[5] - v0[2]};
    double e2[3] = {tv[6] - v0[0], tv[7] - v0[1], tv[8] - v0[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return -1.0;
    double inv = 1.0 / det;
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    return fabs(fmin(fmin(u, v), 1.0 - u - v));
}
