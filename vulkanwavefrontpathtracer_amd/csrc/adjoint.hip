// Path-replay adjoint of the frame loop (Renderer::renderPath,
// Renderer.cpp:175-199, with this library's per-material albedo and
// emission).  The image estimator of one pixel is
//
//     L = sum_pass sum_b  T_b * Y_b,     T_b = prod_{k<b} a[m_k],
//     Y_b = (alpha_b + e[m_b]) * vis_b / spp   (hit at vertex b)
//     Y_b = background / spp                    (miss: vis = 1, path ends)
//
// with m_k the material hit at vertex k, alpha the clamped Lambert term,
// vis the fraction of unblocked shadow rays.  Paths (hits, directions,
// visibility) do not depend on albedo or emission, so under common random
// numbers L is a polynomial in them and
//
//     dL/de[m] = sum over vertices b with m_b = m of T_b * vis_b / spp
//     dL/da[m] = sum over vertices b, k < b with m_k = m of
//                (T_b without the factor a[m_k]) * Y_b / spp-scaled term.
//
// One thread per pixel replays that pixel's passes in order: it regenerates
// the primary ray, re-traces every vertex and shadow ray with the trace
// stage's own traversal, and draws the same random numbers from the same
// per-pixel seed stream as the wavefront stages, so vertices are recomputed
// instead of stored.  The products "without one factor" are kept per earlier
// vertex (E_k), so no division by an albedo is needed.  Gradients are
// accumulated in LDS per workgroup and flushed with one global atomic per
// touched entry.
#include "adjoint.h"
#include "traverse.h"
#include "shading.h"

#include <cfloat>

namespace wfpt {

constexpr int ADJ_BLOCK = 64;

template <bool LDS_GRAD>
__global__ __launch_bounds__(ADJ_BLOCK) void adjoint_kernel(const AdjointParams p) {
    const float4* nodes = p.nodes4;
    __shared__ int s_stack[LDS_DEPTH * ADJ_BLOCK];
    extern __shared__ float s_grad[];  // [3 * nmat] albedo, then [3 * nmat] emission
    const int nm3 = 3 * p.nmat;
    if (LDS_GRAD) {
        for (int k = threadIdx.x; k < 2 * nm3; k += ADJ_BLOCK) s_grad[k] = 0.0f;
        __syncthreads();
    }
    // gradient accumulation into LDS or straight into the outputs; the
    // address space is fixed at compile time (no generic-pointer atomics)
    auto add_ga = [&](int k, float v) {
        if (LDS_GRAD) atomicAdd(&s_grad[k], v);
        else atomicAdd(&p.grad_albedo[k], v);
    };
    auto add_ge = [&](int k, float v) {
        if (LDS_GRAD) atomicAdd(&s_grad[nm3 + k], v);
        else atomicAdd(&p.grad_emission[k], v);
    };

    const int64_t np = (int64_t)p.width * p.height;
    const int64_t i = (int64_t)blockIdx.x * ADJ_BLOCK + threadIdx.x;
    if (i < np) {
        const int pixel = p.index_to_pixel[i];
        uint32_t seed = tea((uint32_t)pixel, (uint32_t)p.frame, 16);
        const float4 g4 = p.d_pixels[pixel];
        const float g[3] = {g4.x, g4.y, g4.z};
        const float spp = (float)p.spp;
        float Lr = 0.f, Lg = 0.f, Lb = 0.f;
        for (int pass = 0; pass < p.spp; ++pass) {
            const int si = pass + p.spp * (p.frame - 1);
            const V3 d0 = primary_direction(p.s2w, pixel, p.width, p.height, halton(si, 2),
                                            halton(si, 3), p.origin[0], p.origin[1], p.origin[2]);
            float4 ro = make_float4(p.origin[0], p.origin[1], p.origin[2], 0.0f);
            float4 rd = make_float4(d0.x, d0.y, d0.z, p.max_dist);
            float T[3] = {1.0f, 1.0f, 1.0f};           // decreases, filled with 1
            float E[ADJ_MAX_VERTS][3];                 // T without the factor of vertex k
            float acc[ADJ_MAX_VERTS][3];               // sum of g * E_k * Y over later vertices
            int mat[ADJ_MAX_VERTS];
            int nv = 0;
#pragma unroll
            for (int k = 0; k < ADJ_MAX_VERTS; ++k) {
                mat[k] = 0;
                for (int c = 0; c < 3; ++c) E[k][c] = acc[k][c] = 0.0f;
            }
            for (int b = 0; b <= p.depth; ++b) {
                WtRay wr;
                const bool live = (rd.w >= ro.w) && wt_setup(ro.x, ro.y, ro.z, rd.x, rd.y, rd.z, wr);
                const TraceOut tr =
                    traverse<true, ADJ_BLOCK, false>(nodes, p.tris, ro, rd, wr, live, s_stack);
                int tri;
                const int4 hit = hit_record(p.tris, wr, tr, &tri);
                float Y[3];
                if (tri < 0) {
                    // reconstructSmooth miss branch; its shadow rays are dead
                    // (tmax = -1), so the visibility is exactly 1.
                    const float aux[3] = {T[0] * p.background[0] / spp, T[1] * p.background[1] / spp,
                                          T[2] * p.background[2] / spp};
                    Lr += aux[0];
                    Lg += aux[1];
                    Lb += aux[2];
                    for (int c = 0; c < 3; ++c) Y[c] = p.background[c] / spp;
#pragma unroll
                    for (int k = 0; k < ADJ_MAX_VERTS; ++k)
                        if (k < nv)
                            for (int c = 0; c < 3; ++c) acc[k][c] += g[c] * E[k][c] * Y[c];
                    break;
                }
                V3 nrm;
                const int m = clamp_material(shading_normal(p.shade, tri, (uint32_t)hit.w, nrm), p.nmat);
                const V3 ff = faceforward3(nrm, v3(rd.x, rd.y, rd.z));
                const float t = __int_as_float(hit.z);
                const V3 x = v3(ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
                const float alpha = smooth_alpha(ff, x, p.light[0], p.light[1], p.light[2]);
                const float e[3] = {p.emission[3 * m], p.emission[3 * m + 1], p.emission[3 * m + 2]};
                const float aux[3] = {T[0] * (alpha + e[0]) / spp, T[1] * (alpha + e[1]) / spp,
                                      T[2] * (alpha + e[2]) / spp};
                // shadow rays of this vertex (raygenShadow + trace any-hit +
                // reconstructShadow), with the seed as it stands
                const V3 off = shadow_offset(seed, pixel);
                const V3 so = shadow_origin(ro, rd, t, false);
                float vis = 0.0f;
                for (int s = 0; s < p.S; ++s) {
                    float4 sa, sb;
                    shadow_ray(s, p.S, off, so, false, p.light[0], p.light[1], p.light[2], p.radius,
                               sa, sb);
                    WtRay swr;
                    const bool slive = (sb.w >= sa.w) && wt_setup(sa.x, sa.y, sa.z, sb.x, sb.y, sb.z, swr);
                    const TraceOut st =
                        traverse<false, ADJ_BLOCK, false>(nodes, p.tris, sa, sb, swr, slive, s_stack);
                    if (st.slot < 0) vis += 1.0f;
                }
                vis /= (float)p.S;
                Lr += aux[0] * vis;
                Lg += aux[1] * vis;
                Lb += aux[2] * vis;
                for (int c = 0; c < 3; ++c) {
                    Y[c] = (alpha + e[c]) * vis / spp;
                    const float de = g[c] * T[c] * vis / spp;
                    if (de != 0.0f) add_ge(3 * m + c, de);
                }
#pragma unroll
                for (int k = 0; k < ADJ_MAX_VERTS; ++k)
                    if (k < nv)
                        for (int c = 0; c < 3; ++c) acc[k][c] += g[c] * E[k][c] * Y[c];
                if (b == p.depth) break;
                // shade/extend (raygenPath.comp:78-170 without roulette)
                const float r0 = randf(seed);
                const float r1 = randf(seed);
                const V3 dn = cosine_direction(r0, r1, ff);
                const float eps = 1.0e-4f;
                ro = make_float4(x.x + dn.x * eps, x.y + dn.y * eps, x.z + dn.z * eps, eps);
                rd = make_float4(dn.x, dn.y, dn.z, FLT_MAX);
                const float w[3] = {p.albedo[3 * m], p.albedo[3 * m + 1], p.albedo[3 * m + 2]};
#pragma unroll
                for (int k = 0; k < ADJ_MAX_VERTS; ++k) {
                    if (k < nv)
                        for (int c = 0; c < 3; ++c) E[k][c] *= w[c];
                    if (k == nv) {
                        mat[k] = m;
                        for (int c = 0; c < 3; ++c) E[k][c] = T[c];
                    }
                }
                ++nv;
                for (int c = 0; c < 3; ++c) T[c] = T[c] * w[c];
            }
#pragma unroll
            for (int k = 0; k < ADJ_MAX_VERTS; ++k)
                if (k < nv)
                    for (int c = 0; c < 3; ++c)
                        if (acc[k][c] != 0.0f) add_ga(3 * mat[k] + c, acc[k][c]);
        }
        if (p.replay) p.replay[pixel] = make_float4(Lr, Lg, Lb, 1.0f);
    }
    if (LDS_GRAD) {
        __syncthreads();
        for (int k = threadIdx.x; k < 2 * nm3; k += ADJ_BLOCK) {
            const float v = s_grad[k];
            if (v != 0.0f) atomicAdd(k < nm3 ? &p.grad_albedo[k] : &p.grad_emission[k - nm3], v);
        }
    }
}

int adjoint_launch(const AdjointParams& p, hipStream_t s) {
    const int64_t np = (int64_t)p.width * p.height;
    if (np == 0) return WF_OK;
    const size_t lds = sizeof(float) * 6 * (size_t)p.nmat;
    // The replay always walks the 4-wide hierarchy: closest hits are exact
    // and any-hit visibility is the same whichever hierarchy the forward
    // pass traversed.
    const unsigned g = grid_for(np, ADJ_BLOCK);
    if (lds <= ADJ_LDS_GRAD_BYTES)
        adjoint_kernel<true><<<g, ADJ_BLOCK, lds, s>>>(p);
    else
        adjoint_kernel<false><<<g, ADJ_BLOCK, 0, s>>>(p);
    return launch_check("adjoint_kernel");
}

}  // namespace wfpt
