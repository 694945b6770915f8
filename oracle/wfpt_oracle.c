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
 *             FP64 one on random and constructed edge cases; the FP32 BVH
 *             trace equal to the FP32 brute force record for record.
 *   RNGs    - TEA, Sobol, Halton and the LCG (the C standard's sample rand())
 *             against published values; the LCG's full low-24-bit period; the
 *             Jenkins mix against its inverse derived from the shader's shifts;
 *             shadow sample points one per stratum.
 *   shading - the Duff et al. basis orthonormal and right-handed; cosine-lobe
 *             moments and distribution over LCG draws; the tone map's
 *             chromaticity and Reinhard curve.
 *   frame   - FP64 frame: empty scene, headlit plane and closed tessellated
 *             sphere closed forms; agreement with the FP32 stage composition
 *             once the replay takes its discrete choices; the FP64 adjoint
 *             against central differences.
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
#pragma omp parallel for schedule(static)
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

/* 10-bit origin coordinates within the scene box. */
static void origin_q10(const float* ray, const float mn[3], const float sz[3], uint32_t q[3]) {
    for (int k = 0; k < 3; ++k) q[k] = (uint32_t)(clamp01f((ray[k] - mn[k]) / sz[k]) * 1023.0f);
}

/* 10-bit-per-axis origin Morton code: bit 3k+2 = x bit k, 3k+1 = y, 3k = z. */
static uint32_t origin30(const float* ray, const float mn[3], const float sz[3]) {
    uint32_t q[3];
    origin_q10(ray, mn, sz, q);
    uint32_t key = 0;
    for (int b = 0; b < 10; ++b)
        for (int k = 0; k < 3; ++k) key |= ((q[k] >> b) & 1u) << (3 * b + (2 - k));
    return key;
}

/* Octahedral map (Cigolle et al. 2014) of the direction, `bits` bits per
 * axis. */
static void oct_q(const float* ray, int bits, uint32_t* qu_out, uint32_t* qv_out) {
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
    *qu_out = (uint32_t)(clamp01f(u) * scale);
    *qv_out = (uint32_t)(clamp01f(v) * scale);
}

/* ... as a 2-D Morton code with u above v. */
static uint32_t oct_code(const float* ray, int bits) {
    uint32_t qu, qv;
    oct_q(ray, bits, &qu, &qv);
    uint32_t code = 0;
    for (int b = 0; b < bits; ++b) code |= (((qu >> b) & 1u) << (2 * b + 1)) | (((qv >> b) & 1u) << (2 * b));
    return code;
}

OR_API void or_keys_origin30(const float* rays, int64_t n, const float mn[3], const float mx[3],
                             uint32_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) keys[i] = origin30(rays + 8 * i, mn, sz);
}

OR_API void or_keys_oct60(const float* rays, int64_t n, const float mn[3], const float mx[3],
                          uint64_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        keys[i] = ((uint64_t)origin30(rays + 8 * i, mn, sz) << 30) | oct_code(rays + 8 * i, 15);
}

/* The same 30 origin and 30 direction bits as one 5-D Morton code (this
 * library's WF_KEY_ORIGIN_DIR_MORTON60): level b = 9..0 holds, from its top,
 * u bit b+5, v bit b+5, x bit b, y bit b, z bit b at key bits 10+5b+4..0;
 * below them the direction's low bits, u bit b at 2b+1 and v bit b at 2b. */
OR_API void or_keys_morton5d60(const float* rays, int64_t n, const float mn[3], const float mx[3],
                               uint64_t* keys) {
    float sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t q[3], qu, qv;
        origin_q10(rays + 8 * i, mn, sz, q);
        oct_q(rays + 8 * i, 15, &qu, &qv);
        uint64_t key = 0;
        for (int b = 0; b < 10; ++b) {
            const int at = 10 + 5 * b;
            key |= (uint64_t)((qu >> (b + 5)) & 1u) << (at + 4);
            key |= (uint64_t)((qv >> (b + 5)) & 1u) << (at + 3);
            key |= (uint64_t)((q[0] >> b) & 1u) << (at + 2);
            key |= (uint64_t)((q[1] >> b) & 1u) << (at + 1);
            key |= (uint64_t)((q[2] >> b) & 1u) << at;
        }
        for (int b = 0; b < 5; ++b)
            key |= ((uint64_t)((qu >> b) & 1u) << (2 * b + 1)) | ((uint64_t)((qv >> b) & 1u) << (2 * b));
        keys[i] = key;
    }
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
#pragma omp parallel for schedule(static)
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

/* ------------------------------------------------------- ray generation */
/* initSeeds.comp:7-19 (TEA) */
static uint32_t tea(uint32_t v0, uint32_t v1, int rounds) {
    uint32_t s0 = 0u;
    for (int n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
OR_API uint32_t or_tea(uint32_t a, uint32_t b, int rounds) { return tea(a, b, rounds); }

OR_API void or_init_seeds(uint32_t* seeds, int64_t n, int frame) {
    for (int64_t i = 0; i < n; ++i) seeds[i] = tea((uint32_t)i, (uint32_t)frame, 16);
}

/* raygenPath.comp:16-26 */
static float randf(uint32_t* seed) {
    *seed = 1103515245u * *seed + 12345u;
    return (float)(*seed & 0x00FFFFFFu) / 16777216.0f;
}
/* n successive draws from seed: the generator states and the floats. */
OR_API void or_lcg(uint32_t seed, int64_t n, uint32_t* states, float* r) {
    for (int64_t i = 0; i < n; ++i) {
        float x = randf(&seed);
        if (states) states[i] = seed;
        if (r) r[i] = x;
    }
}

/* raygenPrimary.comp:11-22 */
static float halton(int i, int base) {
    float r = 0.0f, f = 1.0f / (float)base;
    int j = i;
    while (j > 0) { r += f * (float)(j % base); j /= base; f /= (float)base; }
    return r;
}
OR_API float or_halton(int i, int base) { return halton(i, base); }

/* raygenPrimary.comp:57-84; s2w column-major */
static void primary_dir_f32(const float* m, const float* origin, int pixel, int w, int h, float hx,
                            float hy, float d[3]) {
    float x = (float)(pixel % w) + hx, y = (float)(pixel / w) + hy;
    float nx = 2.0f * x / (float)w - 1.0f, ny = 2.0f * y / (float)h - 1.0f;
    float wx = ((m[0] * nx + m[4] * ny) + m[8] * 0.0f) + m[12] * 1.0f;
    float wy = ((m[1] * nx + m[5] * ny) + m[9] * 0.0f) + m[13] * 1.0f;
    float wz = ((m[2] * nx + m[6] * ny) + m[10] * 0.0f) + m[14] * 1.0f;
    float ww = ((m[3] * nx + m[7] * ny) + m[11] * 0.0f) + m[15] * 1.0f;
    float px = wx / ww, py = wy / ww, pz = wz / ww;
    float dx = px - origin[0], dy = py - origin[1], dz = pz - origin[2];
    float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    d[0] = dx * inv; d[1] = dy * inv; d[2] = dz * inv;
}

OR_API void or_raygen_primary(const float* m, const float* origin, const int32_t* i2p, int w, int h,
                              int sample_index, float max_dist, float* rays) {
    float hx = halton(sample_index, 2), hy = halton(sample_index, 3);
    for (int64_t i = 0; i < (int64_t)w * h; ++i) {
        float* r = rays + 8 * i;
        primary_dir_f32(m, origin, i2p[i], w, h, hx, hy, r + 4);
        r[0] = origin[0]; r[1] = origin[1]; r[2] = origin[2]; r[3] = 0.0f;
        r[7] = max_dist;
    }
}

/* raygenShadow.comp:10-20 */
static void jenkins_mix(uint32_t* a, uint32_t* b, uint32_t* c) {
    *a -= *b; *a -= *c; *a ^= (*c >> 13);
    *b -= *c; *b -= *a; *b ^= (*a << 8);
    *c -= *a; *c -= *b; *c ^= (*b >> 13);
    *a -= *b; *a -= *c; *a ^= (*c >> 12);
    *b -= *c; *b -= *a; *b ^= (*a << 16);
    *c -= *a; *c -= *b; *c ^= (*b >> 5);
    *a -= *b; *a -= *c; *a ^= (*c >> 3);
    *b -= *c; *b -= *a; *b ^= (*a << 10);
    *c -= *a; *c -= *b; *c ^= (*b >> 15);
}
/* raygenShadow.comp:22-40 */
static void sobol2(uint32_t i, float* x, float* y) {
    uint32_t r1 = 0, r2 = 0;
    for (uint32_t v1 = 1u << 31, v2 = 3u << 30; i > 0u; i >>= 1) {
        if (i & 1u) { r1 ^= v1; r2 ^= v2 << 1; }
        v1 |= v1 >> 1;
        v2 ^= v2 >> 1;
    }
    *x = (float)r1 / 4294967296.0f;
    *y = (float)r2 / 4294967296.0f;
}
OR_API void or_sobol2(uint32_t i, float* xy) { sobol2(i, xy, xy + 1); }
OR_API void or_jenkins(uint32_t* abc) { jenkins_mix(abc, abc + 1, abc + 2); }

/* Shadow sample point in [-1,1]^3 for sample i of S with the pixel's
 * Cranley-Patterson offset (raygenShadow.comp:91-118). */
static void shadow_point(uint32_t seed, int pixel, int i, int S, float p[3]) {
    uint32_t ha = seed + (uint32_t)pixel, hb = 0x9e3779b9u, hc = 0x9e3779b9u;
    jenkins_mix(&ha, &hb, &hc);
    jenkins_mix(&ha, &hb, &hc);
    const float two_m32 = 2.3283064365386963e-10f;
    float off[3] = {(float)ha * two_m32, (float)hb * two_m32, (float)hc * two_m32};
    sobol2((uint32_t)i, &p[0], &p[1]);
    p[2] = ((float)i + 0.5f) / (float)S;
    for (int k = 0; k < 3; ++k) {
        p[k] += off[k];
        if (p[k] >= 1.0f) p[k] -= 1.0f;
        p[k] = 2.0f * p[k] - 1.0f;
    }
}

OR_API void or_shadow_point(uint32_t seed, int pixel, int i, int S, float p[3]) {
    shadow_point(seed, pixel, i, S, p);
}

OR_API void or_raygen_shadow(const uint32_t* seeds, const float* in_rays, const int32_t* in_hits,
                             const int32_t* in_s2i, int64_t begin, int64_t end, int S,
                             const float* light, float radius, float* out_rays, int32_t* out_s2i,
                             int32_t* out_i2s) {
    for (int64_t j = 0; j < end - begin; ++j) {
        int64_t src = begin + j;
        int pixel = in_s2i[src];
        uint32_t seed = seeds[pixel];
        const float* r = in_rays + 8 * src;
        int tri = in_hits[4 * src];
        float t;
        memcpy(&t, &in_hits[4 * src + 2], 4);
        float adv = fmaxf(t - 1.0e-5f, 0.0f);
        float o[3] = {r[0] + r[4] * adv, r[1] + r[5] * adv, r[2] + r[6] * adv};
        if (tri == -1) o[0] = o[1] = o[2] = 0.0f;
        for (int i = 0; i < S; ++i) {
            int64_t out = (int64_t)S * j + i;
            float p[3];
            shadow_point(seed, pixel, i, S, p);
            float tg[3] = {light[0] + radius * p[0], light[1] + radius * p[1], light[2] + radius * p[2]};
            float d[3] = {tg[0] - o[0], tg[1] - o[1], tg[2] - o[2]};
            float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            float len = sqrtf(dd);
            float inv = 1.0f / sqrtf(dd);
            float* q = out_rays + 8 * out;
            q[0] = o[0]; q[1] = o[1]; q[2] = o[2]; q[3] = 1.0e-5f;
            if (tri == -1) { q[4] = 1.0f; q[5] = 0.0f; q[6] = 0.0f; q[7] = -1.0f; }
            else { q[4] = d[0] * inv; q[5] = d[1] * inv; q[6] = d[2] * inv; q[7] = len; }
            out_s2i[out] = (int32_t)out;
            out_i2s[out] = (int32_t)out;
        }
    }
}

/* geometrytypes.glsl:18-47 with the packed barycentrics of the hit */
static int shading_normal(const float* tri_n, const int32_t* tri_mat, int64_t tri, uint32_t bary,
                          float n[3]) {
    float bu = (float)(bary & 0xFFFFu) / 65535.0f, bv = (float)(bary >> 16) / 65535.0f;
    float bw = 1.0f - bu - bv;
    const float* a = tri_n + 9 * tri;
    float s[3];
    for (int k = 0; k < 3; ++k) s[k] = a[k] * bw + a[3 + k] * bu + a[6 + k] * bv;
    float inv = 1.0f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int k = 0; k < 3; ++k) n[k] = s[k] * inv;
    return tri_mat[tri];
}

/* raygenPath.comp:28-37: the branchless orthonormal basis of Duff et al.
 * 2017 around the unit normal n (fp32). */
static void duff_basis(const float n[3], float b1[3], float b2[3]) {
    float s = n[2] >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (s + n[2]), b = n[0] * n[1] * a;
    b1[0] = 1.0f + s * n[0] * n[0] * a; b1[1] = s * b; b1[2] = -s * n[0];
    b2[0] = b; b2[1] = s + n[1] * n[1] * a; b2[2] = -n[1];
}
OR_API void or_duff_basis(const float n[3], float b1[3], float b2[3]) { duff_basis(n, b1, b2); }

/* raygenPath.comp:39-49 (fp32; sinf/cosf of the host libm): cosine-weighted
 * direction about n from two uniform draws. */
static void cosine_dir(float r0, float r1, const float n[3], float out[3]) {
    const float two_pi = 2.0f * 3.14159265358979323846f;
    float theta = two_pi * r1, radius = sqrtf(r0);
    float x = radius * sinf(theta), z = radius * cosf(theta), y = sqrtf(1.0f - r0);
    float b1[3], b2[3];
    duff_basis(n, b1, b2);
    for (int k = 0; k < 3; ++k) out[k] = x * b1[k] + z * b2[k] + y * n[k];
}
OR_API void or_cosine_dir(float r0, float r1, const float n[3], float out[3]) {
    cosine_dir(r0, r1, n, out);
}

/* The same basis and direction in fp64 (the fp64 frame oracle's bounce). */
static void duff_basis_f64(const double n[3], double b1[3], double b2[3]) {
    double s = n[2] >= 0.0 ? 1.0 : -1.0;
    double a = -1.0 / (s + n[2]), b = n[0] * n[1] * a;
    b1[0] = 1.0 + s * n[0] * n[0] * a; b1[1] = s * b; b1[2] = -s * n[0];
    b2[0] = b; b2[1] = s + n[1] * n[1] * a; b2[2] = -n[1];
}
static void cosine_dir_f64(double r0, double r1, const double n[3], double out[3]) {
    double theta = 2.0 * 3.14159265358979323846 * r1, rad = sqrt(r0);
    double x = rad * sin(theta), z = rad * cos(theta), y = sqrt(1.0 - r0);
    double b1[3], b2[3];
    duff_basis_f64(n, b1, b2);
    for (int k = 0; k < 3; ++k) out[k] = x * b1[k] + z * b2[k] + y * n[k];
}
/* the same with the basis branch s (+1 or -1) given instead of taken from
 * the sign of n.z */
static void cosine_dir_f64_branch(double r0, double r1, const double n[3], double s,
                                  double out[3]) {
    double theta = 2.0 * 3.14159265358979323846 * r1, rad = sqrt(r0);
    double x = rad * sin(theta), z = rad * cos(theta), y = sqrt(1.0 - r0);
    double a = -1.0 / (s + n[2]), bb = n[0] * n[1] * a;
    double b1[3] = {1.0 + s * n[0] * n[0] * a, s * bb, -s * n[0]};
    double b2[3] = {bb, s + n[1] * n[1] * a, -n[1]};
    for (int k = 0; k < 3; ++k) out[k] = x * b1[k] + z * b2[k] + y * n[k];
}
OR_API void or_duff_basis_f64(const double n[3], double b1[3], double b2[3]) { duff_basis_f64(n, b1, b2); }
OR_API void or_cosine_dir_f64(double r0, double r1, const double n[3], double out[3]) {
    cosine_dir_f64(r0, r1, n, out);
}

/* raygenPath.comp:78-170 with the deterministic (input-order) compaction
 * and pixel-keyed seeds; returns the survivor count.  Not bit-exact against
 * the device (sinf/cosf differ in the last ulps): compared with a
 * tolerance. */
OR_API int64_t or_shade_extend(const int32_t* geom_first, const float* tri_n,
                               const int32_t* tri_mat, const float* albedo, int nmat,
                               uint32_t* seeds, const float* in_rays, const int32_t* in_hits,
                               const int32_t* in_s2i, int64_t n, float* out_rays,
                               int32_t* out_s2i, float* decreases, int rr) {
    int64_t k = 0;
    for (int64_t j = 0; j < n; ++j) {
        const int32_t* h = in_hits + 4 * j;
        if (h[0] < 0) continue;
        int pixel = in_s2i[j];
        uint32_t seed = seeds[pixel];
        const float* r = in_rays + 8 * j;
        int64_t tri = (int64_t)geom_first[h[1]] + h[0];
        float nrm[3];
        int m = shading_normal(tri_n, tri_mat, tri, (uint32_t)h[3], nrm);
        if (m < 0 || m >= nmat) m = 0;
        float dn = nrm[0] * r[4] + nrm[1] * r[5] + nrm[2] * r[6];
        if (!(dn < 0.0f)) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
        float r0 = randf(&seed), r1 = randf(&seed);
        float t;
        memcpy(&t, &h[2], 4);
        float p[3] = {r[0] + r[4] * t, r[1] + r[5] * t, r[2] + r[6] * t};
        float d[3];
        cosine_dir(r0, r1, nrm, d);
        float w[3] = {albedo[3 * m], albedo[3 * m + 1], albedo[3 * m + 2]};
        int alive = 1;
        float* dec = decreases + 4 * (int64_t)pixel;
        if (rr) {
            float q = fmaxf(fmaxf(dec[0] * w[0], dec[1] * w[1]), dec[2] * w[2]);
            q = fminf(fmaxf(q, 0.05f), 1.0f);
            float r2 = randf(&seed);
            if (r2 >= q) alive = 0;
            else { w[0] /= q; w[1] /= q; w[2] /= q; }
        }
        seeds[pixel] = seed;
        if (!alive) continue;
        float* o = out_rays + 8 * k;
        const float eps = 1.0e-4f;
        o[0] = p[0] + d[0] * eps; o[1] = p[1] + d[1] * eps; o[2] = p[2] + d[2] * eps; o[3] = eps;
        o[4] = d[0]; o[5] = d[1]; o[6] = d[2]; o[7] = 3.402823466e+38f;
        out_s2i[k] = pixel;
        dec[0] *= w[0]; dec[1] *= w[1]; dec[2] *= w[2];
        ++k;
    }
    return k;
}

/* ------------------------------------------------------- reconstruction */
/* reconstructSmooth.comp:44-93 (+ emission of the hit material) */
OR_API void or_reconstruct_smooth(const int32_t* geom_first, const float* tri_n,
                                  const int32_t* tri_mat, const float* emission, int nmat,
                                  const float* rays, const int32_t* hits, const int32_t* s2i,
                                  float* pixels, const float* decreases, int64_t n,
                                  const float* light, const float* bg, int spp) {
    float fs = (float)spp;
    for (int64_t i = 0; i < n; ++i) {
        int pixel = s2i[i];
        float* c = pixels + 4 * (int64_t)pixel;
        const float* dec = decreases + 4 * (int64_t)pixel;
        const int32_t* h = hits + 4 * i;
        float col[3] = {c[0], c[1], c[2]};
        if (h[0] == -1) {
            for (int k = 0; k < 3; ++k) col[k] += dec[k] * bg[k] / fs;
        } else {
            const float* r = rays + 8 * i;
            float nrm[3];
            int64_t tri = (int64_t)geom_first[h[1]] + h[0];
            int m = shading_normal(tri_n, tri_mat, tri, (uint32_t)h[3], nrm);
            if (m < 0 || m >= nmat) m = 0;
            float t;
            memcpy(&t, &h[2], 4);
            float p[3] = {r[0] + r[4] * t, r[1] + r[5] * t, r[2] + r[6] * t};
            float L[3] = {light[0] - p[0], light[1] - p[1], light[2] - p[2]};
            float inv = 1.0f / sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
            for (int k = 0; k < 3; ++k) L[k] *= inv;
            float dn = nrm[0] * r[4] + nrm[1] * r[5] + nrm[2] * r[6];
            if (!(dn < 0.0f)) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2]; }
            float alpha = fmaxf(nrm[0] * L[0] + nrm[1] * L[1] + nrm[2] * L[2], 0.1f);
            for (int k = 0; k < 3; ++k) col[k] += dec[k] * (alpha + emission[3 * m + k]) / fs;
        }
        c[0] = col[0]; c[1] = col[1]; c[2] = col[2]; c[3] = 1.0f;
    }
}

/* reconstructShadow.comp:39-59 */
OR_API void or_reconstruct_shadow(const int32_t* shadow_hits, const int32_t* in_s2i,
                                  const int32_t* shadow_i2s, const float* in_px, float* out_px,
                                  int64_t begin, int64_t end, int S, int replace) {
    for (int64_t j = 0; j < end - begin; ++j) {
        int pixel = in_s2i[begin + j];
        const float* c = in_px + 4 * (int64_t)pixel;
        float sh = 0.0f;
        for (int i = 0; i < S; ++i)
            if (shadow_hits[4 * (int64_t)shadow_i2s[j * S + i]] == -1) sh += 1.0f;
        sh /= (float)S;
        float* o = out_px + 4 * (int64_t)pixel;
        float v[4] = {c[0] * sh, c[1] * sh, c[2] * sh, 1.0f};
        if (replace) for (int k = 0; k < 4; ++k) o[k] = v[k];
        else for (int k = 0; k < 4; ++k) o[k] += v[k];
    }
}

/* interpolateColors.comp:10-57.  mat3(a..i) fills columns, so M*v uses the
 * written values column-wise: x = c0*v.x + c3*v.y + c6*v.z. */
static void mat3_cols(const float* c, const float v[3], float o[3]) {
    o[0] = (c[0] * v[0] + c[3] * v[1]) + c[6] * v[2];
    o[1] = (c[1] * v[0] + c[4] * v[1]) + c[7] * v[2];
    o[2] = (c[2] * v[0] + c[5] * v[1]) + c[8] * v[2];
}
static void tonemap(const float in[3], float key, float white, float out[3]) {
    static const float rgb2xyz[9] = {0.5141364f, 0.3238786f, 0.16036376f, 0.265068f, 0.67023428f,
                                     0.06409157f, 0.0241188f, 0.1228178f, 0.84442666f};
    static const float xyz2rgb[9] = {2.5651f, -1.1665f, -0.3986f, -1.0217f, 1.9777f,
                                     0.0439f, 0.0753f, -0.2543f, 1.1892f};
    float XYZ[3];
    mat3_cols(rgb2xyz, in, XYZ);
    float sum = (XYZ[0] + XYZ[1]) + XYZ[2];
    float Y = XYZ[1], x = XYZ[0] / sum, y = XYZ[1] / sum;
    float Lp = Y * key / 0.5f;
    Y = (Lp * (1.0f + Lp / (white * white))) / (1.0f + Lp);
    float X2[3] = {Y * x / y, Y, Y * (1.0f - x - y) / y};
    mat3_cols(xyz2rgb, X2, out);
}
OR_API void or_tonemap(const float* in, float key, float white, float* out) { tonemap(in, key, white, out); }

/* interpolateColors.comp:78-97 */
OR_API void or_interpolate(float* frame, float* pixels, int64_t n, int fi, float key, float white) {
    for (int64_t i = 0; i < n; ++i) {
        float* f = frame + 4 * i;
        float* p = pixels + 4 * i;
        float c[3];
        if (fi > 1) {
            float a = 1.0f / (float)fi, b = a * ((float)fi - 1.0f);
            for (int k = 0; k < 3; ++k) c[k] = a * f[k] + b * p[k];
            p[0] = c[0]; p[1] = c[1]; p[2] = c[2]; p[3] = 1.0f;
        } else {
            for (int k = 0; k < 4; ++k) p[k] = f[k];
            c[0] = f[0]; c[1] = f[1]; c[2] = f[2];
        }
        float t[3];
        tonemap(c, key, white, t);
        f[0] = t[0]; f[1] = t[1]; f[2] = t[2]; f[3] = 1.0f;
    }
}

/* =================================================================
 * FP64 frame oracle: the whole Renderer::renderPath frame
 * (Renderer.cpp:175-199, :552-654) evaluated per pixel in double
 * precision with the same random-number streams (pixel-keyed TEA seed,
 * LCG draws, Sobol/Hammersley shadow points with the Jenkins offset), an
 * fp64 Moller-Trumbore intersector and a median-split BVH built here.
 * Output: the accumulated linear framebuffer (frame 1) as float32 RGBA.
 * ================================================================= */

typedef struct { double lo[3], hi[3]; int left, right, first, count; } bnode;
typedef struct { bnode* nodes; int nn; int* order; const float* tv; } cbvh;

static void tri_bounds(const float* tv, int t, double lo[3], double hi[3]) {
    for (int a = 0; a < 3; ++a) {
        double v0 = tv[9 * t + a], v1 = tv[9 * t + 3 + a], v2 = tv[9 * t + 6 + a];
        lo[a] = fmin(v0, fmin(v1, v2));
        hi[a] = fmax(v0, fmax(v1, v2));
    }
}

static int cb_build(cbvh* b, int first, int count) {
    int id = b->nn++;
    bnode* nd = &b->nodes[id];
    for (int a = 0; a < 3; ++a) { nd->lo[a] = INFINITY; nd->hi[a] = -INFINITY; }
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = first; k < first + count; ++k) {
        double lo[3], hi[3];
        tri_bounds(b->tv, b->order[k], lo, hi);
        for (int a = 0; a < 3; ++a) {
            nd->lo[a] = fmin(nd->lo[a], lo[a]); nd->hi[a] = fmax(nd->hi[a], hi[a]);
            double c = 0.5 * (lo[a] + hi[a]);
            clo[a] = fmin(clo[a], c); chi[a] = fmax(chi[a], c);
        }
    }
    if (count <= 4) { nd->left = nd->right = -1; nd->first = first; nd->count = count; return id; }
    int ax = 0;
    for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
    double mid = 0.5 * (clo[ax] + chi[ax]);
    int i = first, j = first + count - 1;
    while (i <= j) {
        double lo[3], hi[3];
        tri_bounds(b->tv, b->order[i], lo, hi);
        if (0.5 * (lo[ax] + hi[ax]) < mid) ++i;
        else { int t = b->order[i]; b->order[i] = b->order[j]; b->order[j] = t; --j; }
    }
    int nl = i - first;
    if (nl == 0 || nl == count) nl = count / 2;
    int l = cb_build(b, first, nl);
    int r = cb_build(b, first + nl, count - nl);
    nd = &b->nodes[id];
    nd->left = l; nd->right = r; nd->first = 0; nd->count = 0;
    return id;
}

static double mt_hit(const float* tv, const double o[3], const double d[3], double* u_, double* v_) {
    double v0[3] = {tv[0], tv[1], tv[2]};
    double e1[3] = {tv[3] - v0[0], tv[4] - v0[1], tv[5] - v0[2]};
    double e2[3] = {tv[6] - v0[0], tv[7] - v0[1], tv[8] - v0[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return -1.0;
    double inv = 1.0 / det;
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    if (u < 0.0 || u > 1.0) return -1.0;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return -1.0;
    *u_ = u; *v_ = v;
    return (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
}

/* Ambiguity of an fp64 query, used to recognise path vertices where an fp32
 * pipeline may legitimately decide differently.  The fp32 ray is the fp64 ray
 * moved sideways by at most lat(t) = tol0 + tol_eps * t world units at
 * distance t (origin error plus direction error), and the fp32 intersection
 * itself is exact to about EPS_BARY barycentric units.  The query is
 * ambiguous when
 *   - the ray crosses a triangle's plane, before the accepted hit, within
 *     EPS_BARY barycentric units or lat(t) / cos world units (the crossing
 *     point slides by lat / cos on a plane seen at incidence cos) of one of
 *     its edges;
 *   - two hits lie within max(EPS_T * t, lat(t)) of each other (the tie rule);
 *   - a hit lies that close to either end of the interval. */
#define EPS_BARY 1e-5
#define EPS_T 1e-5

typedef struct {
    double u, v, t;
    double cos;    /* |cos| of the angle between the ray and the plane's normal */
    double edge;   /* signed distance (world units, in the plane) to the nearest edge
                      line: >= 0 inside the triangle */
    double n[3];   /* unit normal of the plane */
} line_hit;

/* The ray's line against the triangle's plane, without inside test; returns
 * 0 for a parallel ray. */
static int mt_line(const float* tv, const double o[3], const double d[3], line_hit* h) {
    double v0[3] = {tv[0], tv[1], tv[2]};
    double e1[3] = {tv[3] - v0[0], tv[4] - v0[1], tv[5] - v0[2]};
    double e2[3] = {tv[6] - v0[0], tv[7] - v0[1], tv[8] - v0[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return 0;
    double inv = 1.0 / det;
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    h->u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    h->v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    h->t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2],
           nz = e1[0] * e2[1] - e1[1] * e2[0];
    double area2 = sqrt(nx * nx + ny * ny + nz * nz);
    double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    h->cos = fabs(det) / (area2 * dl);
    /* barycentric coordinate b_i times the height of vertex i over the
     * opposite edge (= 2A / |edge|) is the distance to that edge's line */
    double e3[3] = {e2[0] - e1[0], e2[1] - e1[1], e2[2] - e1[2]};
    double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    double l3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    double w = 1.0 - h->u - h->v;
    h->edge = fmin(fmin(w * area2 / l3, h->u * area2 / l2), h->v * area2 / l1);
    h->n[0] = nx / area2; h->n[1] = ny / area2; h->n[2] = nz / area2;
    return 1;
}

typedef struct {
    double tol0, tol_eps;   /* lateral uncertainty of the fp32 ray: tol0 + tol_eps * t */
    double along;           /* how far the fp32 origin may sit from the exact one in any
                               direction by rounding (of the previous hit distance and the
                               point built from it): it moves the distance to a plane at
                               the interval's start */
    double on[3];           /* unit normal of the surface the ray leaves (0 for a camera
                               ray): the lateral uncertainty runs within that surface, so it
                               moves the distance to a plane by sin(dihedral) times itself */
} ray_tol;

/* closest (any = 0) or any hit in [tmin, tmax]; returns triangle or -1.
 * If amb != NULL, *amb is set when the answer is ambiguous in the sense
 * above (boxes are then expanded so that near misses are examined) and
 * *cos_out receives the incidence of the accepted hit. */
static int cb_trace(const cbvh* b, const double o[3], const double d[3], double tmin, double tmax,
                    int any, double* t_out, double* u_out, double* v_out, int* amb,
                    const ray_tol* tl, double* cos_out) {
    if (!(tmax >= tmin)) return -1;
    int stack[128], sp = 0, node = 0, best = -1;
    double best_t = tmax, bu = 0, bv = 0, bcos = 1;
    double near_t = INFINITY, second_t = INFINITY;
    int certain_any = 0;
    double any_t = 0.0;
    double inv[3];
    const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double t0_ = tl ? tl->tol0 : 0.0, te_ = tl ? tl->tol_eps : 0.0;
    const double along_ = tl ? tl->along : 0.0;
    /* lateral uncertainty (world units) at parameter t, and the same in t units */
#define LAT(t) (t0_ + te_ * fabs(t) * dl)
#define TTOL(t) fmax(EPS_T * fmax(fabs(t), 1.0), LAT(t) / dl)
    for (int a = 0; a < 3; ++a) inv[a] = 1.0 / (fabs(d[a]) > 1e-300 ? d[a] : copysign(1e-300, d[a]));
    stack[sp++] = 0;
    while (sp) {
        node = stack[--sp];
        const bnode* nd = &b->nodes[node];
        double tn = tmin, tf = amb ? best_t + 4 * TTOL(best_t) + 1e-12 : best_t;
        int miss = 0;
        for (int a = 0; a < 3; ++a) {
            double grow = amb ? 1e-4 * (nd->hi[a] - nd->lo[a]) + 1e-9 + 2 * LAT(tf < 1e30 ? tf : 0) : 0.0;
            double t0 = (nd->lo[a] - grow - o[a]) * inv[a], t1 = (nd->hi[a] + grow - o[a]) * inv[a];
            double lo = fmin(t0, t1), hi = fmax(t0, t1);
            tn = fmax(tn, lo - 1e-9 * fabs(lo) - 1e-12);
            tf = fmin(tf, hi + 1e-9 * fabs(hi) + 1e-12);
            if (tn > tf) { miss = 1; break; }
        }
        if (miss) continue;
        if (nd->left < 0) {
            for (int k = nd->first; k < nd->first + nd->count; ++k) {
                int t = b->order[k];
                line_hit lh;
                if (!mt_line(b->tv + 9 * t, o, d, &lh)) continue;
                const double u = lh.u, v = lh.v, th = lh.t;
                double margin = fmin(fmin(u, v), 1.0 - u - v);
                int inside = u >= 0.0 && u <= 1.0 && v >= 0.0 && u + v <= 1.0;
                int at_end = 0;   /* within the fp32 bound of an interval end */
                if (amb) {
                    double tol = TTOL(th);
                    double lat = LAT(th) / fmax(lh.cos, 1e-3);
                    /* the fp32 origin also sits a few ulps of its coordinates
                     * away, by the rounding of the point it was built from
                     * (along, any direction) and anywhere within the previous
                     * hit point's lateral uncertainty (within the surface it
                     * left), which move its distance to this plane, seen
                     * along the ray through it */
                    double mag = fmax(fmax(fabs(o[0]), fabs(o[1])), fabs(o[2])) + 1.0;
                    double cn = tl ? lh.n[0] * tl->on[0] + lh.n[1] * tl->on[1] + lh.n[2] * tl->on[2] : 0.0;
                    double sin_dihedral = sqrt(fmax(0.0, 1.0 - cn * cn));
                    double tol_end = fmax(EPS_T * fabs(th),
                                          (1.2e-7 * mag + along_ + LAT(th) * sin_dihedral) /
                                              fmax(lh.cos, 1e-4) / dl);
                    if ((fabs(margin) < EPS_BARY || fabs(lh.edge) < lat) && th >= tmin - tol &&
                        th <= tmax + tol)
                        near_t = fmin(near_t, th);
                    at_end = fabs(th - tmin) <= tol_end ||
                             (isfinite(tmax) && tmax < 1e30 && fabs(th - tmax) <= tol_end);
                    if (inside && at_end) near_t = fmin(near_t, th);
                }
                if (!inside || !(th >= tmin && th <= best_t)) {
                    if (inside && th >= tmin && th < second_t && best >= 0 && t != best) second_t = fmin(second_t, th);
                    continue;
                }
                if (any) {
                    /* any hit: the interval stays whole, so that a certain
                     * blocker further along settles an uncertain first one */
                    if (best < 0) { best = t; any_t = th; bu = u; bv = v; bcos = lh.cos; }
                    if (!amb || (margin >= EPS_BARY && lh.edge >= LAT(th) / fmax(lh.cos, 1e-3) &&
                                 !at_end)) {
                        certain_any = 1;
                        goto done;
                    }
                    continue;
                }
                if (th == best_t && best >= 0 && t > best) { second_t = fmin(second_t, th); continue; }
                if (best >= 0) second_t = fmin(second_t, best_t);
                best = t; best_t = th; bu = u; bv = v; bcos = lh.cos;
            }
        } else {
            stack[sp++] = nd->right;
            stack[sp++] = nd->left;
        }
    }
done:
    if (amb) {
        double tol = TTOL(best_t);
        if (any) {
            if (!certain_any && near_t <= tmax + tol) *amb = 1;
        } else {
            if (near_t <= best_t + tol) *amb = 1;
            if (best >= 0 && second_t <= best_t + tol) *amb = 1;
        }
    }
#undef LAT
#undef TTOL
    if (best >= 0) { *t_out = any ? any_t : best_t; *u_out = bu; *v_out = bv; }
    if (cos_out) *cos_out = bcos;
    return best;
}

typedef struct {
    int w, h, spp, depth, S, frame, rr, rr_start;
    float light[3], bg[3], origin[3], radius, max_dist, s2w[16];
    int pixel_stride;   /* > 1: render only pixels 0, k, 2k, ... (the others are left as given) */
} or_frame_cfg;

static void dnorm(double v[3]) {
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l; v[1] /= l; v[2] /= l;
}

/* shading normal (exact barycentrics), faceforwarded against d */
static int d_normal(const float* tri_n, const int32_t* tri_mat, int t, double u, double v,
                    const double d[3], double n[3], double* dot_out) {
    const float* a = tri_n + 9 * (int64_t)t;
    double w = 1.0 - u - v;
    for (int k = 0; k < 3; ++k) n[k] = a[k] * w + a[3 + k] * u + a[6 + k] * v;
    dnorm(n);
    const double dn = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    if (dot_out) *dot_out = dn / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(dn < 0.0)) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
    return tri_mat[t];
}

/* Error model of the fp32 pipeline: a position rounded to fp32 after a few
 * operations is within FP32_POS * |p| of the exact one, and a unit direction
 * built from a few fp32 operations (sin/cos of 2*pi*r, sqrt, normalize) is
 * within FP32_DIR radians of the exact one. */
#define FP32_POS (4.0 * 5.960464477539063e-08)
#define FP32_DIR 2e-6

/* lateral uncertainty of an fp32 hit point: the ray's own uncertainty at
 * distance t, slid along the surface at incidence cos, plus rounding of
 * the point */
static double hit_uncertainty(const ray_tol* tl, const double o[3], const double d[3], double t,
                              double cs) {
    double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
    double mag = fabs(p[0]) + fabs(p[1]) + fabs(p[2]) + fabs(t) * dl;
    return (tl->tol0 + tl->tol_eps * fabs(t) * dl) / fmax(cs, 1e-3) + FP32_POS * mag;
}

/* How far the fp32 shading normal of a hit on triangle t may turn from the
 * exact one: the 16-bit barycentrics (packUnorm2x16) that interpolate it, and
 * the hit point's own uncertainty lat (world units in the surface), over
 * which the interpolated normal turns by up to var per smallest triangle
 * height (var: the largest difference of its vertex normals, or of one
 * component of them, comp 0-2; a component equal at the three vertices
 * interpolates exactly). */
static double normal_turn_of(const float* tri_n, const float* tv, int t, double lat, int comp) {
    const float* a = tri_n + 9 * (int64_t)t;
    double var = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                if (comp < 0 || k == comp) var = fmax(var, fabs((double)a[3 * i + k] - a[3 * j + k]));
    if (comp >= 0 && var == 0.0) return 0.0;   /* equal components interpolate exactly */
    const float* p = tv + 9 * (int64_t)t;
    double e[3][3], len[3];
    for (int k = 0; k < 3; ++k) {
        const float* q0 = p + 3 * k;
        const float* q1 = p + 3 * ((k + 1) % 3);
        for (int c = 0; c < 3; ++c) e[k][c] = (double)q1[c] - q0[c];
        len[k] = sqrt(e[k][0] * e[k][0] + e[k][1] * e[k][1] + e[k][2] * e[k][2]);
    }
    double cx = e[0][1] * e[1][2] - e[0][2] * e[1][1], cy = e[0][2] * e[1][0] - e[0][0] * e[1][2],
           cz = e[0][0] * e[1][1] - e[0][1] * e[1][0];
    double area2 = sqrt(cx * cx + cy * cy + cz * cz);
    double height = area2 / fmax(fmax(len[0], len[1]), fmax(len[2], 1e-30));
    return FP32_DIR + 2.0 * var / 65535.0 + (height > 0 ? lat * var / height : var);
}

static double normal_turn(const float* tri_n, const float* tv, int t, double lat) {
    return normal_turn_of(tri_n, tv, t, lat, -1);
}

/* direction uncertainty of a bounce: fp32 arithmetic plus the shading
 * normal's turn, which the basis built on it turns the bounce by at most
 * about three times */
static double bounce_dir_uncertainty(const float* tri_n, const float* tv, int t, double lat) {
    return FP32_DIR + 3.0 * normal_turn(tri_n, tv, t, lat);
}

/* unit geometric normal of triangle t */
static void tri_unit_normal(const float* tv, int t, double n[3]) {
    const float* a = tv + 9 * (int64_t)t;
    double e1[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
    double e2[3] = {a[6] - a[0], a[7] - a[1], a[8] - a[2]};
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
}

/* visibility fraction of the S shadow rays from the hit of (o, d, t) */
/* A decision of the fp32 pipeline (hint) that differs from the fp64 one is
 * followed; *follow records whether the fp64 decision was ambiguous there
 * (bit 0: within the fp32 error bound) or not (bit 1: a real disagreement). */
/* lateral uncertainty (world units) beyond which a replayed path is no
 * longer pinned by the fp64 replay (bit 3 of the follow flags) */
#define OR_LAT_DIVERGED 1e-3

typedef struct {
    uint8_t flags;
    int vertex;                 /* path vertex being decided (set by the caller) */
    int32_t first[4];           /* first unambiguous disagreement: vertex, kind
                                   (1 hit, 2 + sample shadow, -1 roulette), fp64, fp32 */
} follow_rec;
static void note_follow(follow_rec* f, int ambiguous, int kind, int mine, int theirs) {
    f->flags |= ambiguous ? 1 : 2;
    if (!ambiguous && f->first[0] < 0) {
        f->first[0] = f->vertex; f->first[1] = kind; f->first[2] = mine; f->first[3] = theirs;
    }
}

static double d_shadow(const cbvh* b, const or_frame_cfg* c, int hit, const double o[3],
                       const double d[3], double t, uint32_t seed, int pixel, int* amb,
                       double lat_hit, const uint8_t* hint_vis, follow_rec* follow) {
    if (hit < 0) return 1.0;
    double adv = fmax(t - 1.0e-5, 0.0);
    double p[3] = {o[0] + d[0] * adv, o[1] + d[1] * adv, o[2] + d[2] * adv};
    int vis = 0;
    for (int i = 0; i < c->S; ++i) {
        float q[3];
        shadow_point(seed, pixel, i, c->S, q);
        double dir[3];
        for (int k = 0; k < 3; ++k) dir[k] = (double)c->light[k] + (double)c->radius * q[k] - p[k];
        double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int k = 0; k < 3; ++k) dir[k] /= len;
        double tt, uu, vv;
        /* the fp32 shadow ray runs from a point lat_hit away to a light
         * sample rounded to fp32 */
        double lmag = fabs(c->light[0]) + fabs(c->light[1]) + fabs(c->light[2]) + c->radius;
        double pmag = fmax(fmax(fabs(p[0]), fabs(p[1])), fabs(p[2]));
        ray_tol tl = {lat_hit + FP32_POS * lmag, 2 * FP32_DIR, 2.4e-7 * (fabs(t) + pmag), {0, 0, 0}};
        tri_unit_normal(b->tv, hit, tl.on);
        int a = 0;
        int seen = cb_trace(b, p, dir, 1.0e-5, len, 1, &tt, &uu, &vv, &a, &tl, NULL) < 0;
        if (hint_vis && hint_vis[i] <= 1 && hint_vis[i] != seen) {
            note_follow(follow, a, 2 + i, seen, hint_vis[i]);
            seen = hint_vis[i];
        }
        *amb |= a;
        vis += seen;
    }
    return (double)vis / c->S;
}

/* reconstructSmooth of one vertex: out = dec * y with y = (max(N.L, 0.1) +
 * emission[m]) / spp for a hit and background / spp for a miss; returns the
 * material (-1 for a miss) and y. */
static int d_smooth(const cbvh* b, const or_frame_cfg* c, const float* tri_n, const int32_t* tri_mat,
                    const float* emission, int nmat, int hit, const double o[3], const double d[3],
                    double t, double u, double v, const double dec[3], double out[3], double y[3]) {
    (void)b;
    if (hit < 0) {
        for (int k = 0; k < 3; ++k) {
            y[k] = c->bg[k] / c->spp;
            out[k] = dec[k] * y[k];
        }
        return -1;
    }
    double n[3];
    int m = d_normal(tri_n, tri_mat, hit, u, v, d, n, NULL);
    if (m < 0 || m >= nmat) m = 0;
    double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
    double L[3] = {c->light[0] - p[0], c->light[1] - p[1], c->light[2] - p[2]};
    dnorm(L);
    double alpha = fmax(n[0] * L[0] + n[1] * L[1] + n[2] * L[2], 0.1);
    for (int k = 0; k < 3; ++k) {
        y[k] = (alpha + emission[3 * m + k]) / c->spp;
        out[k] = dec[k] * y[k];
    }
    return m;
}

#define OR_MAX_VERTS 16

/* Gradient bookkeeping of one path (the fp64 twin of the adjoint kernel):
 * E[k] = throughput without the albedo factor of extension k, mat[k] its
 * material. */
typedef struct {
    double E[OR_MAX_VERTS][3];
    int mat[OR_MAX_VERTS];
    int nv;
} path_grad;

/* vertex contribution dec * y * vis with loss gradient g: emission gradient
 * of the vertex's material and albedo gradients of the earlier extensions */
static void grad_vertex(const path_grad* pg, int m, const double dec[3], const double y[3],
                        double vis, const double g[3], double spp, double* ga, double* ge) {
    if (!ga) return;
    for (int c = 0; c < 3; ++c) {
        if (m >= 0) ge[3 * m + c] += g[c] * dec[c] * vis / spp;
        for (int k = 0; k < pg->nv; ++k) ga[3 * pg->mat[k] + c] += g[c] * pg->E[k][c] * y[c] * vis;
    }
}

/* Frame 1 of Renderer::renderPath in fp64 (see the ambiguity model above).
 * With d_pixels (w*h float4, rgb used) non-NULL, also the fp64 adjoint:
 * grad_albedo / grad_emission (3*nmat doubles, overwritten) receive
 * sum_pixels d_pixels . dpixel/dparameter. */
/* Take the fp32 pipeline's hit (hint: triangle, or -1 for a miss) where it
 * differs from the fp64 one: the hinted triangle's plane is intersected in
 * fp64 (barycentrics may fall a rounding error outside it). */
static int follow_hit(const cbvh* b, int hit, int hint, int ambiguous, const double o[3],
                      const double d[3], double* t, double* u, double* v, double* cs,
                      follow_rec* follow) {
    if (hint == hit || hint < -1) return hit;
    note_follow(follow, ambiguous, 1, hit, hint);
    if (hint == -1) return -1;
    line_hit lh;
    if (!mt_line(b->tv + 9 * (int64_t)hint, o, d, &lh)) return hit;
    *t = lh.t; *u = lh.u; *v = lh.v; *cs = lh.cos;
    return hint;
}

OR_API void or_render_frame_f64(const float* tv, const float* tri_n, const int32_t* tri_mat,
                                int64_t ntri, const float* albedo, const float* emission, int nmat,
                                const or_frame_cfg* c, float* out_rgba, uint8_t* out_amb,
                                const float* d_pixels, double* grad_albedo, double* grad_emission,
                                const int32_t* hint_tri, const uint8_t* hint_vis,
                                const uint8_t* hint_sign, uint8_t* out_follow,
                                int32_t* out_first) {
    cbvh b;
    b.tv = tv;
    b.nodes = (bnode*)malloc(sizeof(bnode) * (2 * ntri + 1));
    b.order = (int*)malloc(sizeof(int) * (ntri + 1));
    b.nn = 0;
    for (int64_t i = 0; i < ntri; ++i) b.order[i] = (int)i;
    if (ntri > 0) cb_build(&b, 0, (int)ntri);
    const int64_t np = (int64_t)c->w * c->h;
    const int want_grad = d_pixels && grad_albedo && grad_emission;
    if (want_grad)
        for (int k = 0; k < 3 * nmat; ++k) grad_albedo[k] = grad_emission[k] = 0.0;
#pragma omp parallel
    {
        double* ga = want_grad ? (double*)calloc(6 * (size_t)nmat, sizeof(double)) : NULL;
        double* ge = ga ? ga + 3 * nmat : NULL;
        const int64_t stride = c->pixel_stride > 1 ? c->pixel_stride : 1;
        const int64_t nsel = (np + stride - 1) / stride;
#pragma omp for schedule(dynamic, 256)
        for (int64_t sel = 0; sel < nsel; ++sel) {
            const int64_t pixel = sel * stride;
            uint32_t seed = tea((uint32_t)pixel, (uint32_t)c->frame, 16);
            double acc[3] = {0, 0, 0};
            double g[3] = {0, 0, 0};
            if (ga)
                for (int k = 0; k < 3; ++k) g[k] = d_pixels[4 * pixel + k];
            int amb = 0;
            follow_rec follow = {0, 0, {-1, 0, 0, 0}};
            const int V = c->depth + 1;
            /* the fp32 pipeline's decisions of (pass, vertex) for this pixel */
#define HINT_TRI(ps, vx) (hint_tri ? hint_tri[((int64_t)(ps) * V + (vx)) * np + pixel] : -3)
#define HINT_VIS(ps, vx) (hint_vis ? hint_vis + (((int64_t)(ps) * V + (vx)) * np + pixel) * c->S : NULL)
#define HINT_SIGN(ps, vx) (hint_sign ? hint_sign[((int64_t)(ps) * V + (vx)) * np + pixel] : 255)
            for (int pass = 0; pass < c->spp; ++pass) {
                int si = pass + c->spp * (c->frame - 1);
                float hx = halton(si, 2), hy = halton(si, 3);
                /* primary rays are the fp32 ones of raygenPrimary (their near-plane
                 * construction loses ~1e-5 rad to cancellation; the stage is
                 * checked bit for bit on its own): the path is then followed in fp64 */
                float df[3];
                primary_dir_f32(c->s2w, c->origin, (int)pixel, c->w, c->h, hx, hy, df);
                double o[3] = {c->origin[0], c->origin[1], c->origin[2]};
                double d[3] = {df[0], df[1], df[2]};
                double dec[3] = {1, 1, 1};
                double t = 0, u = 0, v = 0, cs = 1;
                ray_tol tl = {0.0, FP32_POS, 0.0, {0, 0, 0}};
                path_grad pg;
                pg.nv = 0;
                int a0 = 0;
                follow.vertex = 0;
                int hit = ntri > 0 ? cb_trace(&b, o, d, 0.0, c->max_dist, 0, &t, &u, &v, &a0, &tl, &cs) : -1;
                hit = follow_hit(&b, hit, HINT_TRI(pass, 0), a0, o, d, &t, &u, &v, &cs, &follow);
                amb |= a0;
                double lat = hit >= 0 ? hit_uncertainty(&tl, o, d, t, cs) : 0.0;
                double aux[3], y[3];
                int m = d_smooth(&b, c, tri_n, tri_mat, emission, nmat, hit, o, d, t, u, v, dec, aux, y);
                double vis = d_shadow(&b, c, hit, o, d, t, seed, (int)pixel, &amb, lat,
                                      HINT_VIS(pass, 0), &follow);
                for (int k = 0; k < 3; ++k) acc[k] += aux[k] * vis;
                grad_vertex(&pg, m, dec, y, vis, g, c->spp, ga, ge);
                for (int bounce = 0; bounce < c->depth; ++bounce) {
                    if (hit < 0) break;
                    double n[3];
                    double ndot = 0.0;
                    int mt = d_normal(tri_n, tri_mat, hit, u, v, d, n, &ndot);
                    /* two sign decisions of the bounce: faceforward (the
                     * normal flipped against the ray) and the basis branch on
                     * the normal's z sign (raygenPath.comp:28-37); within the
                     * fp32 bound either sends the path elsewhere, so the fp32
                     * pipeline's are followed where given, else the pixel is
                     * marked (bit 2) */
                    const double nt = normal_turn(tri_n, b.tv, hit, lat);
                    const double ntz = normal_turn_of(tri_n, b.tv, hit, lat, 2);
                    const int amb_flip = fabs(ndot) <= nt + tl.tol_eps;
                    const int amb_branch = ntz > 0 && fabs(n[2]) <= ntz;
                    double branch = n[2] >= 0.0 ? 1.0 : -1.0;
                    const int hs = HINT_SIGN(pass, bounce);
                    follow.vertex = bounce;
                    if (hs <= 3) {
                        const int flipped = !(ndot < 0.0);
                        if ((hs & 1) != flipped) {
                            note_follow(&follow, amb_flip, 5, flipped, hs & 1);
                            n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
                            branch = n[2] >= 0.0 ? 1.0 : -1.0;
                        }
                        const double want = (hs & 2) ? -1.0 : 1.0;
                        if (want != branch) {
                            note_follow(&follow, amb_branch, 6, branch < 0, want < 0);
                            branch = want;
                        }
                    } else if (amb_flip || amb_branch) {
                        follow.flags |= 4;
                    }
                    if (mt < 0 || mt >= nmat) mt = 0;
                    double r0 = randf(&seed), r1 = randf(&seed);
                    double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
                    double nd[3];
                    cosine_dir_f64_branch(r0, r1, n, branch, nd);
                    double w[3] = {albedo[3 * mt], albedo[3 * mt + 1], albedo[3 * mt + 2]};
                    if (c->rr && bounce >= c->rr_start) {
                        double q = fmax(fmax(dec[0] * w[0], dec[1] * w[1]), dec[2] * w[2]);
                        q = fmin(fmax(q, 0.05), 1.0);
                        double r2 = randf(&seed);
                        const int edge = fabs(r2 - q) < 1e-5;   /* roulette decision on a knife edge */
                        int survive = r2 < q;
                        const int h = HINT_TRI(pass, bounce + 1);
                        follow.vertex = bounce + 1;
                        if (h >= -2 && (h != -2) != survive) {
                            note_follow(&follow, edge, -1, survive, h != -2);
                            survive = h != -2;
                        }
                        amb |= edge;
                        if (!survive) break;
                        for (int k = 0; k < 3; ++k) w[k] /= q;
                    }
                    if (pg.nv < OR_MAX_VERTS) {
                        for (int k = 0; k < pg.nv; ++k)
                            for (int q = 0; q < 3; ++q) pg.E[k][q] *= w[q];
                        for (int q = 0; q < 3; ++q) pg.E[pg.nv][q] = dec[q];
                        pg.mat[pg.nv++] = mt;
                    }
                    for (int k = 0; k < 3; ++k) { dec[k] *= w[k]; o[k] = p[k] + nd[k] * 1.0e-4; d[k] = nd[k]; }
                    tl.tol0 = lat;
                    tl.tol_eps = bounce_dir_uncertainty(tri_n, b.tv, hit, lat);
                    tl.along = 2.4e-7 * (fabs(t) + fmax(fmax(fabs(p[0]), fabs(p[1])), fabs(p[2])));
                    tri_unit_normal(b.tv, hit, tl.on);
                    int ab = 0;
                    follow.vertex = bounce + 1;
                    hit = cb_trace(&b, o, d, 1.0e-4, 3.402823466e+38, 0, &t, &u, &v, &ab, &tl, &cs);
                    hit = follow_hit(&b, hit, HINT_TRI(pass, bounce + 1), ab, o, d, &t, &u, &v, &cs,
                                     &follow);
                    amb |= ab;
                    lat = hit >= 0 ? hit_uncertainty(&tl, o, d, t, cs) : 0.0;
                    /* past this bound the fp32 path and the fp64 one part ways
                     * continuously (each bounce off a curved, interpolated
                     * normal multiplies the deviation), whatever the choices */
                    if (lat > OR_LAT_DIVERGED) follow.flags |= 8;
                    m = d_smooth(&b, c, tri_n, tri_mat, emission, nmat, hit, o, d, t, u, v, dec, aux, y);
                    vis = d_shadow(&b, c, hit, o, d, t, seed, (int)pixel, &amb, lat,
                                   HINT_VIS(pass, bounce + 1), &follow);
                    for (int k = 0; k < 3; ++k) acc[k] += aux[k] * vis;
                    grad_vertex(&pg, m, dec, y, vis, g, c->spp, ga, ge);
                }
            }
            float* px = out_rgba + 4 * pixel;
            px[0] = (float)acc[0]; px[1] = (float)acc[1]; px[2] = (float)acc[2]; px[3] = 1.0f;
            if (out_amb) out_amb[pixel] = (uint8_t)amb;
            if (out_follow) out_follow[pixel] = follow.flags;
            if (out_first) memcpy(out_first + 4 * pixel, follow.first, sizeof follow.first);
#undef HINT_TRI
#undef HINT_VIS
#undef HINT_SIGN
        }
        if (ga) {
#pragma omp critical
            for (int k = 0; k < 3 * nmat; ++k) {
                grad_albedo[k] += ga[k];
                grad_emission[k] += ge[k];
            }
            free(ga);
        }
    }
    free(b.nodes);
    free(b.order);
}

/* fp64 BVH trace for large scenes (closest or any), same output layout as
 * or_trace_f64. */
OR_API void or_trace_bvh_f64(const float* rays, int64_t n, const float* tri_verts, int64_t ntri,
                             int closest, double* out_t, int32_t* out_tri) {
    cbvh b;
    b.tv = tri_verts;
    b.nodes = (bnode*)malloc(sizeof(bnode) * (2 * ntri + 1));
    b.order = (int*)malloc(sizeof(int) * (ntri + 1));
    b.nn = 0;
    for (int64_t i = 0; i < ntri; ++i) b.order[i] = (int)i;
    if (ntri > 0) cb_build(&b, 0, (int)ntri);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
        double t = -1, u, v;
        int hit = ntri > 0 ? cb_trace(&b, o, d, r[3], r[7], !closest, &t, &u, &v, NULL, NULL, NULL) : -1;
        out_t[i] = hit >= 0 ? t : -1.0;
        out_tri[i] = hit;
    }
    free(b.nodes);
    free(b.order);
}

/* fp32 watertight trace over the median-split BVH, for scenes too large for
 * the brute-force scan of or_trace_f32 (trace.comp:27-88 semantics): each
 * candidate triangle goes through the same wt_intersect and the same
 * interval and tie rules, so the records equal or_trace_f32's exactly,
 * provided no box rejects a triangle the fp32 test would accept.  The boxes
 * are therefore widened by 1e-5 of the coordinate magnitudes in play (the
 * fp32 test's effective triangle and t deviate from the exact ones by a few
 * ulps, ~5e-7 relative) and the slab interval is tested in fp64 with slack.
 * Any hit: the smallest triangle index hit within the interval, which is
 * what the brute-force scan returns.  Output layout as or_trace_f32. */
typedef struct { cbvh b; int64_t ntri; } or_bvh;

/* Build once, trace many times (the CPU baseline times the trace alone). */
OR_API or_bvh* or_bvh_create(const float* tri_verts, int64_t ntri) {
    or_bvh* h = (or_bvh*)malloc(sizeof(or_bvh));
    h->ntri = ntri;
    h->b.tv = tri_verts;
    h->b.nodes = (bnode*)malloc(sizeof(bnode) * (2 * ntri + 1));
    h->b.order = (int*)malloc(sizeof(int) * (ntri + 1));
    h->b.nn = 0;
    for (int64_t i = 0; i < ntri; ++i) h->b.order[i] = (int)i;
    if (ntri > 0) cb_build(&h->b, 0, (int)ntri);
    return h;
}

OR_API void or_bvh_destroy(or_bvh* h) {
    if (!h) return;
    free(h->b.nodes);
    free(h->b.order);
    free(h);
}

OR_API void or_bvh_trace_f32(const or_bvh* h, const float* rays, int64_t n, int closest,
                             int32_t* out) {
    const cbvh b = h->b;
    const float* tri_verts = b.tv;
    const int64_t ntri = h->ntri;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float* ry = rays + 8 * i;
        const float tmin = ry[3], tmax = ry[7];
        int32_t best = -1;
        float best_t = tmax, bv = 0, bw = 0, bd = 1;
        wt_ray w;
        if (ntri > 0 && tmax >= tmin && wt_setup(ry, ry + 4, &w)) {
            const double o[3] = {ry[0], ry[1], ry[2]};
            const double d[3] = {ry[4], ry[5], ry[6]};
            double inv[3];
            for (int a = 0; a < 3; ++a)
                inv[a] = 1.0 / (fabs(d[a]) > 1e-300 ? d[a] : copysign(1e-300, d[a]));
            const double omag = fmax(fmax(fabs(o[0]), fabs(o[1])), fabs(o[2]));
            int stack[128], sp = 0;
            stack[sp++] = 0;
            while (sp) {
                const bnode* nd = &b.nodes[stack[--sp]];
                double m = omag;
                for (int a = 0; a < 3; ++a) m = fmax(m, fmax(fabs(nd->lo[a]), fabs(nd->hi[a])));
                const double grow = 1e-5 * (m + 1.0);
                double tn = (double)tmin, tf = closest ? (double)best_t : (double)tmax;
                tn -= 1e-5 * (fabs(tn) + 1.0);
                tf += 1e-5 * (fabs(tf) + 1.0);
                int miss = 0;
                for (int a = 0; a < 3; ++a) {
                    double t0 = (nd->lo[a] - grow - o[a]) * inv[a], t1 = (nd->hi[a] + grow - o[a]) * inv[a];
                    tn = fmax(tn, fmin(t0, t1));
                    tf = fmin(tf, fmax(t0, t1));
                    if (tn > tf) { miss = 1; break; }
                }
                if (miss) continue;
                if (nd->left >= 0) {
                    if (sp + 2 > 128) abort();
                    stack[sp++] = nd->right;
                    stack[sp++] = nd->left;
                    continue;
                }
                for (int k = nd->first; k < nd->first + nd->count; ++k) {
                    const int t_ = b.order[k];
                    const float* tv = tri_verts + 9 * (int64_t)t_;
                    float t, U, V, W, det;
                    if (!wt_intersect(&w, tv, tv + 3, tv + 6, &t, &U, &V, &W, &det)) continue;
                    if (closest) {
                        if (!(t >= tmin && t <= best_t)) continue;
                        if (t == best_t && best >= 0 && t_ > best) continue;
                    } else {
                        if (!(t >= tmin && t <= tmax)) continue;
                        if (best >= 0 && t_ > best) continue;
                    }
                    best = t_; best_t = t; bv = V; bw = W; bd = det;
                }
            }
        }
        int32_t* q = out + 4 * i;
        if (best >= 0) {
            q[0] = best;
            memcpy(&q[1], &best_t, 4);
            q[2] = (int32_t)pack_unorm2x16(bv / bd, bw / bd);
        } else {
            const float m1 = -1.0f;
            q[0] = -1;
            memcpy(&q[1], &m1, 4);
            q[2] = 0;
        }
        q[3] = 0;
    }
}

OR_API void or_trace_bvh_f32(const float* rays, int64_t n, const float* tri_verts, int64_t ntri,
                             int closest, int32_t* out) {
    or_bvh* h = or_bvh_create(tri_verts, ntri);
    or_bvh_trace_f32(h, rays, n, closest, out);
    or_bvh_destroy(h);
}
