// LBVH build: replaces the driver BLAS/TLAS of the reference
// (projects/WaveFrontPathTracer/WaveFrontPathTracer.cpp:202-350 builds one
// BLAS with a geometry per glTF primitive, :355-457 one identity TLAS
// instance; Tracer::setAccererationStructure binds it, Tracer/Tracer.cpp:
// 166-183).
//
// Build (all on the GPU, stream ordered):
//   1. tri_prep   : per-triangle AABB and centroid, scene / centroid bounds
//                   (wave reductions + one ordered-int atomic per wave),
//                   shading records in original order.
//   2. references: a triangle whose box is larger than 1/64 of the scene
//      diagonal (bvh.split_denom) is split (Ernst & Greiner's early split clipping: recursive
//      midpoint splits of the clipped polygon, up to 2^split_depth pieces),
//      so that large triangles (walls, floors) do not inflate the boxes near
//      the root; small triangles are one reference each.
//   3. centroid Morton codes of the references, 21 bits per axis (63-bit
//      keys), sorted with this library's radix sort (reference payload).
//   4. the binary tree over the sorted references (bvh.builder): top-down
//      binned SAH (default; the Morton order only seeds the reference
//      order), or Karras 2012's radix tree over the sorted codes (equal codes
//      split by index); parent links.
//   5. bottom-up AABB refit (second child to arrive at a node merges).
//   6. emit 64-byte BVH2 nodes holding both child boxes; subtrees of at most
//      BVH_LEAF_MAX triangles collapse into one leaf (contiguous slots);
//      child boxes are enlarged by 1e-5 x scene diagonal so that fp32 slab
//      tests never cull a box whose triangles the intersection test hits.
#include "scene.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace wfpt {

// Monotone map float -> uint for atomicMin/atomicMax on floats.
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// bounds[0..5]: scene min xyz (keys), scene max xyz; [6..11] centroid min/max.
__global__ __launch_bounds__(256) void tri_prep_kernel(
    const char* __restrict__ verts, int64_t nverts, int stride, const uint32_t* __restrict__ idx,
    int64_t ntri, const int32_t* __restrict__ geom_first, int ngeom,
    const int32_t* __restrict__ tri_mat, float4* __restrict__ tri_box /* 2 per tri */,
    float4* __restrict__ shade, int32_t* __restrict__ tri_geom, uint32_t* __restrict__ bounds,
    uint32_t* __restrict__ bad_index) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = t < ntri;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float c[3] = {0, 0, 0};
    if (ok) {
        float p[3][3], nrm[3][3];
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = idx[3 * t + k];
            if ((int64_t)vi >= nverts) {
                atomicOr(bad_index, 1u);
                vi = 0;
            }
            const float* v = reinterpret_cast<const float*>(verts + (int64_t)vi * stride);
            for (int a = 0; a < 3; ++a) {
                p[k][a] = v[a];
                nrm[k][a] = v[3 + a];
            }
        }
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(fminf(p[0][a], p[1][a]), p[2][a]);
            hi[a] = fmaxf(fmaxf(p[0][a], p[1][a]), p[2][a]);
            c[a] = (lo[a] + hi[a]) * 0.5f;
        }
        tri_box[2 * t] = make_float4(lo[0], lo[1], lo[2], 0.f);
        tri_box[2 * t + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
        int mat = tri_mat ? tri_mat[t] : 0;
        shade[3 * t + 0] = make_float4(nrm[0][0], nrm[0][1], nrm[0][2], __int_as_float(mat));
        shade[3 * t + 1] = make_float4(nrm[1][0], nrm[1][1], nrm[1][2], 0.f);
        shade[3 * t + 2] = make_float4(nrm[2][0], nrm[2][1], nrm[2][2], 0.f);
        // geometry of this triangle: last g with geom_first[g] <= t
        int lo_g = 0, hi_g = ngeom - 1;
        while (lo_g < hi_g) {
            int mid = (lo_g + hi_g + 1) >> 1;
            if (geom_first[mid] <= t) lo_g = mid;
            else hi_g = mid - 1;
        }
        tri_geom[t] = lo_g;
    }
    float cl[3], ch[3];
    for (int a = 0; a < 3; ++a) {
        cl[a] = ok ? c[a] : INFINITY;
        ch[a] = ok ? c[a] : -INFINITY;
    }
    for (int a = 0; a < 3; ++a) {
        float v0 = wave_min(lo[a]), v1 = wave_max(hi[a]), v2 = wave_min(cl[a]), v3 = wave_max(ch[a]);
        if (lane_id() == 0) {
            if (v0 <= v1) {
                atomicMin(&bounds[a], fkey(v0));
                atomicMax(&bounds[3 + a], fkey(v1));
            }
            if (v2 <= v3) {
                atomicMin(&bounds[6 + a], fkey(v2));
                atomicMax(&bounds[9 + a], fkey(v3));
            }
        }
    }
}

// Sutherland-Hodgman clip of polygon `in` (n vertices) to the half space
// x[axis] <= mid (below) or >= mid (above); returns the vertex count.
constexpr int SPLIT_MAXV = 12;
constexpr int SPLIT_MAX_DEPTH = 12;  // largest bvh.split_depth
__device__ int clip_half(const float (*in)[3], int n, float (*out)[3], int axis, float mid,
                         bool below) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const float* a = in[i];
        const float* b = in[(i + 1) % n];
        float da = below ? (mid - a[axis]) : (a[axis] - mid);
        float db = below ? (mid - b[axis]) : (b[axis] - mid);
        if (da >= 0.f && m < SPLIT_MAXV) {
            out[m][0] = a[0];
            out[m][1] = a[1];
            out[m][2] = a[2];
            ++m;
        }
        if ((da >= 0.f) != (db >= 0.f) && m < SPLIT_MAXV) {
            float t = da / (da - db);
            for (int k = 0; k < 3; ++k) out[m][k] = a[k] + t * (b[k] - a[k]);
            out[m][axis] = mid;
            ++m;
        }
    }
    return m;
}

// Counts (EMIT = false) or writes (EMIT = true) the references of each
// triangle: pieces of the triangle clipped by recursive midpoint splits of
// the largest axis until the piece's extent is <= limit or depth runs out.
template <bool EMIT>
__global__ __launch_bounds__(128) void split_kernel(
    const char* __restrict__ verts, int stride, const uint32_t* __restrict__ idx,
    const float4* __restrict__ tri_box, int64_t ntri, const uint32_t* __restrict__ bounds,
    float frac, int maxd, uint32_t* __restrict__ counts, const int64_t* __restrict__ offsets,
    float4* __restrict__ item_box, uint32_t* __restrict__ item_tri) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    float sx = funkey(bounds[3]) - funkey(bounds[0]);
    float sy = funkey(bounds[4]) - funkey(bounds[1]);
    float sz = funkey(bounds[5]) - funkey(bounds[2]);
    const float limit = frac * sqrtf(sx * sx + sy * sy + sz * sz);
    float4 blo = tri_box[2 * t], bhi = tri_box[2 * t + 1];
    float ext = fmaxf(fmaxf(bhi.x - blo.x, bhi.y - blo.y), bhi.z - blo.z);
    int64_t out = EMIT ? offsets[t] : 0;
    if (ext <= limit || maxd == 0) {
        if (EMIT) {
            item_box[2 * out] = blo;
            item_box[2 * out + 1] = bhi;
            item_tri[out] = (uint32_t)t;
        } else {
            counts[t] = 1;
        }
        return;
    }
    // depth-first stack of pieces: a split replaces the top piece by its two
    // halves, so at most one piece per depth level plus one is pending
    float poly[SPLIT_MAX_DEPTH + 1][SPLIT_MAXV][3];
    int nv[SPLIT_MAX_DEPTH + 1], dep[SPLIT_MAX_DEPTH + 1];
    int sp = 0;
    for (int k = 0; k < 3; ++k) {
        const float* v = reinterpret_cast<const float*>(verts + (int64_t)idx[3 * t + k] * stride);
        poly[0][k][0] = v[0];
        poly[0][k][1] = v[1];
        poly[0][k][2] = v[2];
    }
    nv[0] = 3;
    dep[0] = 0;
    sp = 1;
    uint32_t count = 0;
    while (sp > 0) {
        --sp;
        const int n = nv[sp], d = dep[sp];
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], poly[sp][i][k]);
                hi[k] = fmaxf(hi[k], poly[sp][i][k]);
            }
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
        if (hi[axis] - lo[axis] <= limit || d >= maxd || sp + 2 > SPLIT_MAX_DEPTH + 1) {
            if (EMIT) {
                item_box[2 * out] = make_float4(lo[0], lo[1], lo[2], 0.f);
                item_box[2 * out + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
                item_tri[out] = (uint32_t)t;
                ++out;
            }
            ++count;
            continue;
        }
        const float mid = 0.5f * (lo[axis] + hi[axis]);
        float cur[SPLIT_MAXV][3];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) cur[i][k] = poly[sp][i][k];
        int na = clip_half(cur, n, poly[sp], axis, mid, false);
        if (na >= 3) {
            nv[sp] = na;
            dep[sp] = d + 1;
            ++sp;
        }
        int nb = clip_half(cur, n, poly[sp], axis, mid, true);
        if (nb >= 3) {
            nv[sp] = nb;
            dep[sp] = d + 1;
            ++sp;
        }
    }
    if (!EMIT) counts[t] = count;
}

__device__ __forceinline__ uint64_t spread3_21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x001f00000000ffffull;
    v = (v | (v << 16)) & 0x001f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Morton code of each reference's box centre within the scene bounds.
__global__ __launch_bounds__(256) void centroid_keys_kernel(const float4* __restrict__ item_box,
                                                            int64_t nitem,
                                                            const uint32_t* __restrict__ bounds,
                                                            uint64_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nitem) return;
    float4 lo = item_box[2 * t], hi = item_box[2 * t + 1];
    float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    uint64_t q[3];
    for (int a = 0; a < 3; ++a) {
        float mn = funkey(bounds[a]), mx = funkey(bounds[3 + a]);
        float ext = mx - mn;
        float u = ext > 0.f ? (c[a] - mn) / ext : 0.f;
        u = fminf(fmaxf(u, 0.f), 1.f);
        q[a] = (uint64_t)(u * 2097151.0f);
    }
    keys[t] = (spread3_21(q[0]) << 2) | (spread3_21(q[1]) << 1) | spread3_21(q[2]);
    vals[t] = (uint32_t)t;
}

__device__ __forceinline__ int lbvh_delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll(a ^ b);
    return 64 + __clz(i ^ j);
}

// Internal nodes 0..n-2; combined index space: internal i -> i, leaf s -> n-1+s.
__global__ __launch_bounds__(256) void karras_kernel(const uint64_t* __restrict__ k, int n,
                                                     int2* __restrict__ child,
                                                     int2* __restrict__ range,
                                                     int32_t* __restrict__ parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (lbvh_delta(k, n, i, i + 1) - lbvh_delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = lbvh_delta(k, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lbvh_delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = lbvh_delta(k, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (lbvh_delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int first = min(i, j), last = max(i, j);
    int left = (first == gamma) ? (n - 1 + gamma) : gamma;
    int right = (last == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(first, last);
    parent[left] = i;
    parent[right] = i;
    if (i == 0) parent[0] = -1;
}

// ------------------------------------------------------------ binned SAH
// Top-down binned SAH build (Wald 2007's binning of reference centroids,
// 32 bins per axis, cost = area(left) * count(left) + area(right) *
// count(right)), the alternative to the radix tree (bvh.builder = 1). Nodes
// of more than SAH_SMALL references are split a level at a time, every node
// of the level at once: their references are cut into chunks of SAH_CHUNK
// positions, one workgroup per chunk, for the centroid bounds, the bins and
// the stable partition; one wave per node picks the split. Smaller nodes
// finish their whole subtree in one wave, the references in LDS. Nodes at
// depth SAH_MEDIAN_DEPTH and below, and nodes whose centroids share one bin
// on every axis, split at the median position. The references start in
// Morton order and every partition is stable, so each node's references stay
// in Morton order.
// The build writes the radix tree's child / range / parent arrays and leaves
// the final reference order in the sorted-item array, so refit and emission
// are shared. Numbering: every gap between consecutive positions g, g+1 is
// split by exactly one internal node, which takes index g, except that the
// root and the node of gap 0 trade indices so that the root is 0. The
// numbering is therefore a function of the tree, whatever order the levels'
// nodes are processed in.
constexpr int SAH_BINS = 32;  // a power of two up to the wave width
constexpr int SAH_SMALL = 512;
constexpr int SAH_CHUNK = 4096;
constexpr int SAH_MEDIAN_DEPTH = 40;
constexpr int SAH_BIN_WORDS = 3 * SAH_BINS * 7;  // per axis and bin: count, box lo xyz, box hi xyz
constexpr int SAH_WAVES = 4;                      // small-node waves per workgroup

struct SahNode {          // a node of more than SAH_SMALL references
    int first, count;     // positions [first, first + count)
    int parent, side;     // parent index and child slot (-1 for the root)
    int depth, axis;      // axis of the split, -1: median position
    int split, lc;        // last bin of the left side; references on the left
    uint32_t cmin[3], cmax[3];  // centroid bounds (ordered keys)
    int pad[2];
};
static_assert(sizeof(SahNode) == 64, "SahNode is 64 bytes");

__device__ __forceinline__ float3 ref_centre(const float4* __restrict__ item_box, uint32_t it) {
    const float4 lo = item_box[2 * (int64_t)it], hi = item_box[2 * (int64_t)it + 1];
    return make_float3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
}

__device__ __forceinline__ float f3(const float3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Bin of a centroid coordinate between the node's centroid bounds (the same
// arithmetic in every kernel, so binning and partition agree).
__device__ __forceinline__ int sah_bin(float c, float mn, float mx) {
    const float ext = mx - mn;
    if (!(ext > 0.f)) return 0;
    const int b = (int)((c - mn) / ext * (float)SAH_BINS);
    return min(max(b, 0), SAH_BINS - 1);
}

__device__ __forceinline__ int sah_index(int g, int groot) { return g == groot ? 0 : (g == 0 ? groot : g); }

__device__ __forceinline__ float half_area(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return dx * dy + dy * dz + dz * dx;
}

// The best split over the bins of one node (global or LDS words), evaluated by
// a whole wave, one lane per bin (lanes i and i + SAH_BINS mirror each other
// when the bins are fewer than the lanes). Returns the axis (-1 if
// no plane leaves references on both sides), the last bin of the left side
// and the left count. Ties go to the lower axis, then the lower bin.
__device__ void sah_best_split(const uint32_t* bins, int& axis_out, int& split_out, int& lc_out) {
    const int k = threadIdx.x & (SAH_BINS - 1);
    float best = INFINITY;
    int axis = -1, split = 0, lc = 0;
    for (int a = 0; a < 3; ++a) {
        const uint32_t* w = bins + (a * SAH_BINS + k) * 7;
        uint32_t cnt = w[0];
        float lo[3], hi[3];
        for (int j = 0; j < 3; ++j) {
            lo[j] = cnt ? funkey(w[1 + j]) : INFINITY;
            hi[j] = cnt ? funkey(w[4 + j]) : -INFINITY;
        }
        // inclusive prefix (bins 0..k) and suffix (bins k..SAH_BINS - 1)
        uint32_t pc = cnt, sc = cnt;
        float pl[3], ph[3], sl[3], sh[3];
        for (int j = 0; j < 3; ++j) {
            pl[j] = sl[j] = lo[j];
            ph[j] = sh[j] = hi[j];
        }
        for (int o = 1; o < SAH_BINS; o <<= 1) {
            const uint32_t uc = __shfl_up(pc, o, SAH_BINS), dc = __shfl_down(sc, o, SAH_BINS);
            float ul[3], uh[3], dl[3], dh[3];
            for (int j = 0; j < 3; ++j) {
                ul[j] = __shfl_up(pl[j], o, SAH_BINS);
                uh[j] = __shfl_up(ph[j], o, SAH_BINS);
                dl[j] = __shfl_down(sl[j], o, SAH_BINS);
                dh[j] = __shfl_down(sh[j], o, SAH_BINS);
            }
            if (k >= o) {
                pc += uc;
                for (int j = 0; j < 3; ++j) {
                    pl[j] = fminf(pl[j], ul[j]);
                    ph[j] = fmaxf(ph[j], uh[j]);
                }
            }
            if (k + o < SAH_BINS) {
                sc += dc;
                for (int j = 0; j < 3; ++j) {
                    sl[j] = fminf(sl[j], dl[j]);
                    sh[j] = fmaxf(sh[j], dh[j]);
                }
            }
        }
        // split after bin k: left = prefix(k), right = suffix(k + 1)
        const uint32_t rc = __shfl_down(sc, 1, SAH_BINS);
        float rl[3], rh[3];
        for (int j = 0; j < 3; ++j) {
            rl[j] = __shfl_down(sl[j], 1, SAH_BINS);
            rh[j] = __shfl_down(sh[j], 1, SAH_BINS);
        }
        float cost = INFINITY;
        if (k < SAH_BINS - 1 && pc > 0 && rc > 0)
            cost = half_area(pl[0], pl[1], pl[2], ph[0], ph[1], ph[2]) * (float)pc +
                   half_area(rl[0], rl[1], rl[2], rh[0], rh[1], rh[2]) * (float)rc;
        // argmin over the planes, lowest bin on ties
        float c = cost;
        int b = k;
        uint32_t l = pc;
        for (int o = SAH_BINS / 2; o >= 1; o >>= 1) {
            const float oc = __shfl_xor(c, o, SAH_BINS);
            const int ob = __shfl_xor(b, o, SAH_BINS);
            const uint32_t ol = __shfl_xor(l, o, SAH_BINS);
            if (oc < c || (oc == c && ob < b)) {
                c = oc;
                b = ob;
                l = ol;
            }
        }
        if (c < best) {
            best = c;
            axis = a;
            split = b;
            lc = (int)l;
        }
    }
    axis_out = axis;
    split_out = split;
    lc_out = lc;
}

__device__ __forceinline__ bool sah_goes_left(int axis, int split, int lc, int local, float3 c,
                                              const float* cmn, const float* cmx) {
    if (axis < 0) return local < lc;
    return sah_bin(f3(c, axis), cmn[axis], cmx[axis]) <= split;
}

// Links of a node that has just been split at gap g: its own index, its slot
// in its parent, its range, and its one-reference children (leaves).
__device__ __forceinline__ int sah_link(int first, int count, int lc, int par, int side, int groot,
                                        int n, int2* __restrict__ child, int2* __restrict__ range,
                                        int32_t* __restrict__ parent) {
    const int g = first + lc - 1;
    const int id = par < 0 ? 0 : sah_index(g, groot);
    if (par >= 0) reinterpret_cast<int*>(child)[2 * par + side] = id;
    parent[id] = par;
    range[id] = make_int2(first, first + count - 1);
    if (lc == 1) {
        reinterpret_cast<int*>(child)[2 * id] = n - 1 + first;
        parent[n - 1 + first] = id;
    }
    if (count - lc == 1) {
        reinterpret_cast<int*>(child)[2 * id + 1] = n - 1 + first + lc;
        parent[n - 1 + first + lc] = id;
    }
    return id;
}

__global__ __launch_bounds__(256) void sah_bins_init_kernel(uint32_t* __restrict__ bins, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= words) return;
    const int w = (int)(i % 7);
    bins[i] = w == 0 ? 0u : (w <= 3 ? 0xffffffffu : 0u);
}

// chunk = {node, first position, count, 0}
__global__ __launch_bounds__(256) void sah_cbounds_kernel(const int4* __restrict__ chunks,
                                                          const uint32_t* __restrict__ perm,
                                                          const float4* __restrict__ item_box,
                                                          SahNode* __restrict__ nodes) {
    const int4 ch = chunks[blockIdx.x];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < ch.z; i += 256) {
        const float3 c = ref_centre(item_box, perm[ch.y + i]);
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], f3(c, a));
            mx[a] = fmaxf(mx[a], f3(c, a));
        }
    }
    for (int a = 0; a < 3; ++a) {
        mn[a] = wave_min(mn[a]);
        mx[a] = wave_max(mx[a]);
    }
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) {
            atomicMin(&nodes[ch.x].cmin[a], fkey(mn[a]));
            atomicMax(&nodes[ch.x].cmax[a], fkey(mx[a]));
        }
    }
}

__global__ __launch_bounds__(256) void sah_bin_kernel(const int4* __restrict__ chunks,
                                                      const uint32_t* __restrict__ perm,
                                                      const float4* __restrict__ item_box,
                                                      const SahNode* __restrict__ nodes,
                                                      uint32_t* __restrict__ bins) {
    __shared__ uint32_t s[SAH_BIN_WORDS];
    for (int k = threadIdx.x; k < SAH_BIN_WORDS; k += 256) {
        const int w = k % 7;
        s[k] = w == 0 ? 0u : (w <= 3 ? 0xffffffffu : 0u);
    }
    __syncthreads();
    const int4 ch = chunks[blockIdx.x];
    const SahNode& nd = nodes[ch.x];
    float cmn[3], cmx[3];
    for (int a = 0; a < 3; ++a) {
        cmn[a] = funkey(nd.cmin[a]);
        cmx[a] = funkey(nd.cmax[a]);
    }
    for (int i = threadIdx.x; i < ch.z; i += 256) {
        const uint32_t it = perm[ch.y + i];
        const float4 lo = item_box[2 * (int64_t)it], hi = item_box[2 * (int64_t)it + 1];
        const float3 c = make_float3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
        const uint32_t kl[3] = {fkey(lo.x), fkey(lo.y), fkey(lo.z)};
        const uint32_t kh[3] = {fkey(hi.x), fkey(hi.y), fkey(hi.z)};
        for (int a = 0; a < 3; ++a) {
            uint32_t* w = s + (a * SAH_BINS + sah_bin(f3(c, a), cmn[a], cmx[a])) * 7;
            atomicAdd(w, 1u);
            for (int j = 0; j < 3; ++j) {
                atomicMin(w + 1 + j, kl[j]);
                atomicMax(w + 4 + j, kh[j]);
            }
        }
    }
    __syncthreads();
    uint32_t* g = bins + (int64_t)ch.x * SAH_BIN_WORDS;
    for (int k = threadIdx.x; k < 3 * SAH_BINS; k += 256) {
        const uint32_t* w = s + 7 * k;
        if (w[0] == 0) continue;
        atomicAdd(g + 7 * k, w[0]);
        for (int j = 1; j <= 3; ++j) atomicMin(g + 7 * k + j, w[j]);
        for (int j = 4; j <= 6; ++j) atomicMax(g + 7 * k + j, w[j]);
    }
}

// One wave per node of the level: the split, the links, the children.
// counters: [0] nodes of the next level, [1] small nodes, [2] the root's gap.
__global__ __launch_bounds__(256) void sah_split_kernel(SahNode* __restrict__ nodes, int nl,
                                                        const uint32_t* __restrict__ bins, int n,
                                                        SahNode* __restrict__ next,
                                                        int4* __restrict__ small,
                                                        uint32_t* __restrict__ counters,
                                                        int2* __restrict__ child,
                                                        int2* __restrict__ range,
                                                        int32_t* __restrict__ parent) {
    const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (j >= nl) return;
    const SahNode nd = nodes[j];
    int axis, split, lc;
    sah_best_split(bins + (int64_t)j * SAH_BIN_WORDS, axis, split, lc);
    if (axis < 0 || nd.depth >= SAH_MEDIAN_DEPTH) {
        axis = -1;
        split = 0;
        lc = nd.count / 2;
    }
    if ((threadIdx.x & 63) != 0) return;
    nodes[j].axis = axis;
    nodes[j].split = split;
    nodes[j].lc = lc;
    const int groot = nd.parent < 0 ? nd.first + lc - 1 : (int)counters[2];
    if (nd.parent < 0) counters[2] = (uint32_t)groot;
    const int id = sah_link(nd.first, nd.count, lc, nd.parent, nd.side, groot, n, child, range, parent);
    for (int sd = 0; sd < 2; ++sd) {
        const int f = sd ? nd.first + lc : nd.first;
        const int c = sd ? nd.count - lc : lc;
        if (c > SAH_SMALL) {
            SahNode ch;
            ch.first = f;
            ch.count = c;
            ch.parent = id;
            ch.side = sd;
            ch.depth = nd.depth + 1;
            ch.axis = -1;
            ch.split = 0;
            ch.lc = 0;
            for (int a = 0; a < 3; ++a) {
                ch.cmin[a] = 0xffffffffu;
                ch.cmax[a] = 0u;
            }
            ch.pad[0] = ch.pad[1] = 0;
            next[atomicAdd(&counters[0], 1u)] = ch;
        } else if (c >= 2) {
            small[atomicAdd(&counters[1], 1u)] = make_int4(f, c, id, sd | ((nd.depth + 1) << 1));
        }
    }
}

// Lefts per chunk (the same test as the scatter).
__global__ __launch_bounds__(256) void sah_count_kernel(const int4* __restrict__ chunks,
                                                        const uint32_t* __restrict__ perm,
                                                        const float4* __restrict__ item_box,
                                                        const SahNode* __restrict__ nodes,
                                                        uint32_t* __restrict__ lefts) {
    __shared__ uint32_t s_w[4];
    const int4 ch = chunks[blockIdx.x];
    const SahNode& nd = nodes[ch.x];
    float cmn[3], cmx[3];
    for (int a = 0; a < 3; ++a) {
        cmn[a] = funkey(nd.cmin[a]);
        cmx[a] = funkey(nd.cmax[a]);
    }
    uint32_t cnt = 0;
    for (int i = threadIdx.x; i < ch.z; i += 256) {
        const float3 c = ref_centre(item_box, perm[ch.y + i]);
        cnt += sah_goes_left(nd.axis, nd.split, nd.lc, ch.y + i - nd.first, c, cmn, cmx) ? 1u : 0u;
    }
    for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) lefts[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Per node: left and right references before each of its chunks.
__global__ __launch_bounds__(256) void sah_offsets_kernel(const int* __restrict__ chunk0, int nl,
                                                          const uint32_t* __restrict__ lefts,
                                                          int2* __restrict__ offs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    int l = 0, seen = 0;
    for (int c = chunk0[j]; c < chunk0[j + 1]; ++c) {
        offs[c] = make_int2(l, seen - l);
        l += (int)lefts[c];
        seen += SAH_CHUNK;
    }
}

// Stable partition of each chunk's references into the node's two sides.
__global__ __launch_bounds__(256) void sah_scatter_kernel(const int4* __restrict__ chunks,
                                                          const uint32_t* __restrict__ perm,
                                                          const float4* __restrict__ item_box,
                                                          const SahNode* __restrict__ nodes,
                                                          const int2* __restrict__ offs,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t s_w[4];
    const int4 ch = chunks[blockIdx.x];
    const int cnt = __builtin_amdgcn_readfirstlane(ch.z);  // uniform trip count for the barriers
    const SahNode& nd = nodes[ch.x];
    float cmn[3], cmx[3];
    for (int a = 0; a < 3; ++a) {
        cmn[a] = funkey(nd.cmin[a]);
        cmx[a] = funkey(nd.cmax[a]);
    }
    const int2 o = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int lbase = 0;
    for (int r = 0; r < cnt; r += 256) {
        const int i = r + threadIdx.x;
        const bool valid = i < cnt;
        uint32_t it = 0;
        bool left = false;
        if (valid) {
            it = perm[ch.y + i];
            left = sah_goes_left(nd.axis, nd.split, nd.lc, ch.y + i - nd.first,
                                 ref_centre(item_box, it), cmn, cmx);
        }
        const uint64_t m = __ballot(left);
        if (lane == 0) s_w[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        int before = lbase;
        for (int k = 0; k < wv; ++k) before += (int)s_w[k];
        const int total = (int)(s_w[0] + s_w[1] + s_w[2] + s_w[3]);
        __syncthreads();
        if (valid) {
            const int lb = before + (int)__popcll(m & below);  // lefts before i in this chunk
            const int dst = left ? nd.first + o.x + lb : nd.first + nd.lc + o.y + (i - lb);
            if (dst >= nd.first && dst < nd.first + nd.count) out[dst] = it;
        }
        lbase += total;
    }
}

__global__ __launch_bounds__(256) void sah_copy_kernel(const int4* __restrict__ chunks,
                                                       const uint32_t* __restrict__ src,
                                                       uint32_t* __restrict__ dst) {
    const int4 ch = chunks[blockIdx.x];
    for (int i = threadIdx.x; i < ch.z; i += 256) dst[ch.y + i] = src[ch.y + i];
}

// One wave per small node (task = {first, count, parent, side | depth << 1}):
// the node's whole subtree, depth first, with its references in LDS.
__global__ __launch_bounds__(256) void sah_small_kernel(const int4* __restrict__ tasks, int ntask,
                                                        const float4* __restrict__ item_box,
                                                        uint32_t* __restrict__ perm, int n,
                                                        const uint32_t* __restrict__ counters,
                                                        int2* __restrict__ child,
                                                        int2* __restrict__ range,
                                                        int32_t* __restrict__ parent) {
    __shared__ uint32_t s_perm[SAH_WAVES][SAH_SMALL];
    __shared__ uint32_t s_tmp[SAH_WAVES][SAH_SMALL];
    __shared__ uint32_t s_bins[SAH_WAVES][SAH_BIN_WORDS];
    __shared__ int4 s_stack[SAH_WAVES][SAH_SMALL / 2];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int task = blockIdx.x * SAH_WAVES + wv;
    if (task >= ntask) return;  // whole waves; no workgroup barriers below
    const int4 t = tasks[task];
    const int base = t.x;
    uint32_t* sp_perm = s_perm[wv];
    uint32_t* sp_tmp = s_tmp[wv];
    uint32_t* sp_bins = s_bins[wv];
    int4* sp_stack = s_stack[wv];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = lane; i < t.y; i += 64) sp_perm[i] = perm[base + i];
    int groot = t.z < 0 ? -1 : (int)counters[2];
    int sp = 0;
    if (lane == 0) sp_stack[0] = make_int4(0, t.y, t.z, t.w);
    sp = 1;
    while (sp > 0) {
        __builtin_amdgcn_wave_barrier();
        const int4 e = sp_stack[--sp];
        const int f = e.x, c = e.y, par = e.z, side = e.w & 1, depth = e.w >> 1;
        float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = lane; i < c; i += 64) {
            const float3 cc = ref_centre(item_box, sp_perm[f + i]);
            for (int a = 0; a < 3; ++a) {
                cmn[a] = fminf(cmn[a], f3(cc, a));
                cmx[a] = fmaxf(cmx[a], f3(cc, a));
            }
        }
        for (int a = 0; a < 3; ++a) {
            cmn[a] = wave_min(cmn[a]);
            cmx[a] = wave_max(cmx[a]);
        }
        for (int k = lane; k < SAH_BIN_WORDS; k += 64) {
            const int w = k % 7;
            sp_bins[k] = w == 0 ? 0u : (w <= 3 ? 0xffffffffu : 0u);
        }
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < c; i += 64) {
            const uint32_t it = sp_perm[f + i];
            const float4 lo = item_box[2 * (int64_t)it], hi = item_box[2 * (int64_t)it + 1];
            const float3 cc = make_float3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
            const uint32_t kl[3] = {fkey(lo.x), fkey(lo.y), fkey(lo.z)};
            const uint32_t kh[3] = {fkey(hi.x), fkey(hi.y), fkey(hi.z)};
            for (int a = 0; a < 3; ++a) {
                uint32_t* w = sp_bins + (a * SAH_BINS + sah_bin(f3(cc, a), cmn[a], cmx[a])) * 7;
                atomicAdd(w, 1u);
                for (int j = 0; j < 3; ++j) {
                    atomicMin(w + 1 + j, kl[j]);
                    atomicMax(w + 4 + j, kh[j]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        int axis, split, lc;
        sah_best_split(sp_bins, axis, split, lc);
        if (axis < 0 || depth >= SAH_MEDIAN_DEPTH) {
            axis = -1;
            split = 0;
            lc = c / 2;
        }
        int lbase = 0, rbase = 0;
        for (int i0 = 0; i0 < c; i0 += 64) {
            const int i = i0 + lane;
            const bool valid = i < c;
            uint32_t it = 0;
            bool left = false;
            if (valid) {
                it = sp_perm[f + i];
                left = sah_goes_left(axis, split, lc, i, ref_centre(item_box, it), cmn, cmx);
            }
            const uint64_t ml = __ballot(left), mv = __ballot(valid);
            if (valid) {
                const int lb = (int)__popcll(ml & below), rb = (int)__popcll(mv & ~ml & below);
                sp_tmp[f + (left ? lbase + lb : lc + rbase + rb)] = it;
            }
            lbase += (int)__popcll(ml);
            rbase += (int)__popcll(mv & ~ml);
        }
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < c; i += 64) sp_perm[f + i] = sp_tmp[f + i];
        const int first = base + f;
        if (par < 0) groot = first + lc - 1;
        int id = 0;
        if (lane == 0) id = sah_link(first, c, lc, par, side, groot, n, child, range, parent);
        id = __shfl(id, 0, 64);
        if (lane == 0) {
            if (c - lc >= 2) sp_stack[sp] = make_int4(f + lc, c - lc, id, 1 | ((depth + 1) << 1));
            if (lc >= 2) sp_stack[sp + (c - lc >= 2 ? 1 : 0)] = make_int4(f, lc, id, (depth + 1) << 1);
        }
        sp += (c - lc >= 2 ? 1 : 0) + (lc >= 2 ? 1 : 0);
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < t.y; i += 64) perm[base + i] = sp_perm[i];
}

__global__ __launch_bounds__(256) void refit_kernel(const float4* __restrict__ item_box,
                                                    const uint32_t* __restrict__ sorted_item, int n,
                                                    const int2* __restrict__ child,
                                                    const int32_t* __restrict__ parent,
                                                    float4* __restrict__ box /* 2 per node */,
                                                    uint32_t* __restrict__ flags,
                                                    uint32_t* __restrict__ max_depth) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t t = sorted_item[s];
    int leaf = n - 1 + s;
    float4 lo = item_box[2 * (int64_t)t], hi = item_box[2 * (int64_t)t + 1];
    box[2 * leaf] = lo;
    box[2 * leaf + 1] = hi;
    // depth of this leaf (number of internal ancestors)
    uint32_t depth = 0;
    for (int p = parent[leaf]; p >= 0; p = parent[p]) ++depth;
    atomicMax(max_depth, depth);
    int p = parent[leaf];
    while (p >= 0) {
        __threadfence();
        uint32_t old = atomicAdd(&flags[p], 1u);
        if (old == 0) return;  // the sibling subtree finishes this node
        __threadfence();
        int2 c = child[p];
        float4 al = box[2 * c.x], ah = box[2 * c.x + 1];
        float4 bl = box[2 * c.y], bh = box[2 * c.y + 1];
        box[2 * p] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.f);
        box[2 * p + 1] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.f);
        p = parent[p];
    }
}

__device__ __forceinline__ void child_ref(int c, int n, const int2* __restrict__ range,
                                          int leaf_max, int& ref) {
    if (c >= n - 1) {
        ref = leaf_ref(c - (n - 1), 1);
    } else {
        int2 r = range[c];
        int size = r.y - r.x + 1;
        ref = size <= leaf_max ? leaf_ref(r.x, size) : c;
    }
}

__global__ __launch_bounds__(256) void emit_nodes_kernel(int n, int leaf_max,
                                                         const int2* __restrict__ child,
                                                         const int2* __restrict__ range,
                                                         const float4* __restrict__ box,
                                                         const uint32_t* __restrict__ bounds,
                                                         float4* __restrict__ nodes,
                                                         uint32_t* __restrict__ leaves) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    float dx = funkey(bounds[3]) - funkey(bounds[0]);
    float dy = funkey(bounds[4]) - funkey(bounds[1]);
    float dz = funkey(bounds[5]) - funkey(bounds[2]);
    float margin = 1e-5f * sqrtf(dx * dx + dy * dy + dz * dz) + 1e-30f;
    int2 c = child[i];
    int r0, r1;
    child_ref(c.x, n, range, leaf_max, r0);
    child_ref(c.y, n, range, leaf_max, r1);
    float4 l0 = box[2 * c.x], h0 = box[2 * c.x + 1], l1 = box[2 * c.y], h1 = box[2 * c.y + 1];
    l0.x -= margin; l0.y -= margin; l0.z -= margin; h0.x += margin; h0.y += margin; h0.z += margin;
    l1.x -= margin; l1.y -= margin; l1.z -= margin; h1.x += margin; h1.y += margin; h1.z += margin;
    float4* o = nodes + 4 * (int64_t)i;
    o[0] = make_float4(l0.x, l0.y, l0.z, h0.x);
    o[1] = make_float4(h0.y, h0.z, l1.x, l1.y);
    o[2] = make_float4(l1.z, h1.x, h1.y, h1.z);
    o[3] = make_float4(__int_as_float(r0), __int_as_float(r1), 0.f, 0.f);
    if (r0 < 0) atomicAdd(leaves, 1u);
    if (r1 < 0) atomicAdd(leaves, 1u);
}

// Single-triangle scene: one node with the leaf and an empty box.
__global__ void emit_single_kernel(const float4* __restrict__ tri_box, float4* __restrict__ nodes) {
    float4 lo = tri_box[0], hi = tri_box[1];
    float m = 1e-5f * sqrtf((hi.x - lo.x) * (hi.x - lo.x) + (hi.y - lo.y) * (hi.y - lo.y) +
                            (hi.z - lo.z) * (hi.z - lo.z)) + 1e-30f;
    nodes[0] = make_float4(lo.x - m, lo.y - m, lo.z - m, hi.x + m);
    nodes[1] = make_float4(hi.y + m, hi.z + m, INFINITY, INFINITY);
    nodes[2] = make_float4(INFINITY, -INFINITY, -INFINITY, -INFINITY);
    nodes[3] = make_float4(__int_as_float(leaf_ref(0, 1)), __int_as_float(leaf_ref(0, 1)), 0.f, 0.f);
}

__global__ __launch_bounds__(256) void emit_tris_kernel(const char* __restrict__ verts, int stride,
                                                        const uint32_t* __restrict__ idx,
                                                        const uint32_t* __restrict__ sorted_item,
                                                        const uint32_t* __restrict__ item_tri,
                                                        const int32_t* __restrict__ tri_geom,
                                                        const int32_t* __restrict__ geom_first,
                                                        int64_t nslot, float4* __restrict__ tris) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    uint32_t t = item_tri ? item_tri[sorted_item[s]] : sorted_item[s];
    int g = tri_geom[t];
    int prim = (int)t - geom_first[g];
    float w[3] = {__int_as_float((int)t), __int_as_float(prim), __int_as_float(g)};
    for (int k = 0; k < 3; ++k) {
        const float* v = reinterpret_cast<const float*>(verts + (int64_t)idx[3 * (int64_t)t + k] * stride);
        tris[3 * s + k] = make_float4(v[0], v[1], v[2], w[k]);
    }
}

// ------------------------------------------------------- 4-wide hierarchy
// Every interior node of the 2-wide hierarchy at even depth becomes a 4-wide
// node whose children are its grandchildren (a leaf child is kept as it is).
// A 2-wide node is real (emitted) when it is the root or holds more than
// leaf_max references; its depth counts its real ancestors, all of which are
// real since ranges nest.
__global__ __launch_bounds__(256) void wide_mark_kernel(int n, int leaf_max,
                                                        const int2* __restrict__ range,
                                                        const int32_t* __restrict__ parent,
                                                        uint32_t* __restrict__ flag,
                                                        uint32_t* __restrict__ depth_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int2 r = range[i];
    const bool real = i == 0 || (r.y - r.x + 1) > leaf_max;
    uint32_t depth = 0;
    for (int p = parent[i]; p >= 0; p = parent[p]) ++depth;
    flag[i] = (real && (depth & 1u) == 0u) ? 1u : 0u;
    if (real) atomicMax(depth_out, depth);
}

// Exclusive prefix of 0/1 flags in index order (decoupled look-back, one
// element per thread); the last tile writes the total.
__global__ __launch_bounds__(256) void flag_scan_kernel(const uint32_t* __restrict__ flag, int64_t n,
                                                        uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ tile_counter,
                                                        uint32_t* __restrict__ status,
                                                        uint32_t* __restrict__ total) {
    __shared__ uint32_t s_tile, s_excl, s_wave[4];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t ntiles = (n + 255) / 256;
    const int64_t i = tile * 256 + tid;
    const bool f = i < n && flag[i] != 0u;
    const uint64_t bal = ballot64(f);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t wbase = 0, sum = 0;
    for (int w = 0; w < 4; ++w) {
        if (w < (int)wave) wbase += s_wave[w];
        sum += s_wave[w];
    }
    if (tid == 0) {
        uint32_t excl = 0;
        if (tile == 0) {
            lb_store(&status[0], LB_FLAG_INC | sum);
        } else {
            lb_store(&status[tile], LB_FLAG_AGG | sum);
            for (int64_t k = tile - 1;; --k) {
                uint32_t v;
                do {
                    v = lb_load(&status[k]);
                } while ((v & ~LB_MASK) == 0u);
                excl += v & LB_MASK;
                if (v & LB_FLAG_INC) break;
            }
            lb_store(&status[tile], LB_FLAG_INC | (excl + sum));
        }
        s_excl = excl;
        if (tile == ntiles - 1) *total = excl + sum;
    }
    __syncthreads();
    if (i < n) out[i] = s_excl + wbase + popc_below(bal);
}

__device__ __forceinline__ int wide_exponent(float extent, int e_min) {
    int e = e_min;
    // extent / 2^e <= QMAX - 3: the origin sits one quantum below the low
    // plane and the planes round outwards by one more
    const float x = extent / (float)(NODE4_QMAX - 3);
    if (x > 0.0f) {
        int ex = 0;
        (void)frexpf(x, &ex);  // x = m * 2^ex, m in [0.5, 1): 2^ex >= x
        e = max(e, ex);
    }
    return min(max(e, -126), 127);
}

// Quantized 4-wide node from a flagged 2-wide node.  e_min keeps every
// quantum at least 2^-19 of the larger of the scene's diagonal and its
// largest coordinate magnitude, far above the rounding of the slab test
// (a few ulps, 2^-22 of the same), so the outward rounding by one quantum
// keeps the boxes conservative.
__device__ __forceinline__ int wide_min_exponent(const uint32_t* __restrict__ bounds) {
    float m = 0.0f, d2 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float lo = funkey(bounds[a]), hi = funkey(bounds[3 + a]);
        m = fmaxf(m, fmaxf(fabsf(lo), fabsf(hi)));
        d2 += (hi - lo) * (hi - lo);
    }
    const float size = fmaxf(fmaxf(m, sqrtf(d2)), 1e-30f);
    int ex = 0;
    (void)frexpf(size, &ex);
    return ex - 18;  // quantum >= 2^-18 * 2^ex > 2^-18 * size... >= 2^-19 * size
}

// Two plane values (integers 0..NODE4_QMAX, exact in fp16) as one word of
// fp16 halves, the first in the low half.
__device__ __forceinline__ float half_pair(float a, float b) {
    const uint32_t lo = __builtin_bit_cast(unsigned short, (_Float16)a);
    const uint32_t hi = __builtin_bit_cast(unsigned short, (_Float16)b);
    return __uint_as_float(lo | hi << 16);
}

// Writes one 4-wide node (layout in scene.h) from per-child plane values.
__device__ __forceinline__ void write_node4(float4* __restrict__ o, const float (&origin)[3],
                                            uint32_t meta, const float (&ql)[4][3],
                                            const float (&qh)[4][3], const int (&ref)[4]) {
    o[0] = make_float4(origin[0], origin[1], origin[2], __uint_as_float(meta));
    for (int a = 0; a < 3; ++a)
        o[1 + a] = make_float4(half_pair(ql[0][a], ql[1][a]), half_pair(ql[2][a], ql[3][a]),
                               half_pair(qh[0][a], qh[1][a]), half_pair(qh[2][a], qh[3][a]));
    o[4] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]),
                       __int_as_float(ref[3]));
}

// Quantized planes of one child box in the node frame (origin, scale):
// rounded outwards by one quantum and clamped to the grid.
__device__ __forceinline__ void quantize_child(const float* lo, const float* hi, const float* origin,
                                               const float* scale, float* ql, float* qh) {
    for (int a = 0; a < 3; ++a) {
        ql[a] = fminf(fmaxf(floorf((lo[a] - origin[a]) / scale[a]) - 1.0f, 0.0f), (float)NODE4_QMAX);
        qh[a] = fminf(fmaxf(ceilf((hi[a] - origin[a]) / scale[a]) + 1.0f, 0.0f), (float)NODE4_QMAX);
    }
}

// Empty child slot k: an inverted box (low plane QMAX above high plane 0 on
// every axis), which no slab test with a nonzero scale meets, and a reference
// that is harmless if a degenerate ray meets it anyway (leaf slot 0: one
// extra test of a scene triangle changes no closest hit), so the traversal
// needs no child count.
__device__ __forceinline__ void empty_child(float (&ql)[4][3], float (&qh)[4][3], int (&ref)[4], int k) {
    for (int a = 0; a < 3; ++a) {
        ql[k][a] = (float)NODE4_QMAX;
        qh[k][a] = 0.0f;
    }
    ref[k] = leaf_ref(0, 1);
}

// Single-reference scene: one 4-wide node with the leaf as its only child.
__global__ void wide_single_kernel(const float4* __restrict__ nodes, const uint32_t* __restrict__ bounds,
                                   float4* __restrict__ nodes4) {
    const float4 a = nodes[0], b = nodes[1];
    const float lo[3] = {a.x, a.y, a.z}, hi[3] = {a.w, b.x, b.y};
    const int e_min = wide_min_exponent(bounds);
    float origin[3], scale[3], ql[4][3], qh[4][3];
    int ref[4];
    uint32_t meta = 1u << 24;
    for (int k = 0; k < 3; ++k) {
        const int e = wide_exponent(hi[k] - lo[k], e_min);
        scale[k] = ldexpf(1.0f, e);
        origin[k] = lo[k] - scale[k];
        meta |= (uint32_t)(e + 127) << (8 * k);
    }
    quantize_child(lo, hi, origin, scale, ql[0], qh[0]);
    ref[0] = leaf_ref(0, 1);
    for (int k = 1; k < 4; ++k) empty_child(ql, qh, ref, k);
    write_node4(nodes4, origin, meta, ql, qh, ref);
}

// Child c (0/1) of 2-wide node `node`: its box and reference.
__device__ __forceinline__ void node_child(const float4* __restrict__ nodes, int node, int c,
                                           float* l, float* h, int& r) {
    const float4* nd = nodes + 4 * (int64_t)node;
    if (c == 0) {
        const float4 a = nd[0], b = nd[1], d = nd[3];
        l[0] = a.x; l[1] = a.y; l[2] = a.z; h[0] = a.w; h[1] = b.x; h[2] = b.y;
        r = __float_as_int(d.x);
    } else {
        const float4 b = nd[1], cc = nd[2], d = nd[3];
        l[0] = b.z; l[1] = b.w; l[2] = cc.x; h[0] = cc.y; h[1] = cc.z; h[2] = cc.w;
        r = __float_as_int(d.y);
    }
}

// Children (2-4) of the 4-wide node made from 2-wide node i, as 2-wide
// references (interior node index or leaf).  GREEDY: open the interior child
// with the largest box surface area until there are four children or only
// leaves (the usual collapse of a binary hierarchy into a wide one, e.g.
// Wald, Benthin & Boulos, "Getting Rid of Packets", 2008); otherwise open
// every interior child once (the grandchildren).
template <bool GREEDY>
__device__ __forceinline__ int collapse4(const float4* __restrict__ nodes, int i, float (&lo)[4][3],
                                         float (&hi)[4][3], int (&ref)[4]) {
    int nc = 0;
    if (GREEDY) {
        node_child(nodes, i, 0, lo[0], hi[0], ref[0]);
        node_child(nodes, i, 1, lo[1], hi[1], ref[1]);
        nc = 2;
        while (nc < 4) {
            int best = -1;
            float ba = 0.f;
            for (int k = 0; k < nc; ++k) {
                if (ref[k] < 0) continue;
                const float dx = hi[k][0] - lo[k][0], dy = hi[k][1] - lo[k][1], dz = hi[k][2] - lo[k][2];
                const float a = dx * dy + dy * dz + dz * dx;
                if (best < 0 || a > ba) {
                    ba = a;
                    best = k;
                }
            }
            if (best < 0) break;
            const int r = ref[best];
            node_child(nodes, r, 1, lo[nc], hi[nc], ref[nc]);
            node_child(nodes, r, 0, lo[best], hi[best], ref[best]);
            ++nc;
        }
        return nc;
    }
    for (int c = 0; c < 2; ++c) {
        float l[3], h[3];
        int r;
        node_child(nodes, i, c, l, h, r);
        if (r >= 0) {
            for (int g = 0; g < 2; ++g) {
                node_child(nodes, r, g, lo[nc], hi[nc], ref[nc]);
                ++nc;
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                lo[nc][a] = l[a];
                hi[nc][a] = h[a];
            }
            ref[nc++] = r;
        }
    }
    return nc;
}

// One level of the top-down marking of the greedy 4-wide nodes: every node
// of the frontier flags its interior children and queues them.  need = the
// traversal stack bound along the path (a visit pushes at most nc - 1).
__global__ __launch_bounds__(256) void wide_frontier_kernel(const float4* __restrict__ nodes,
                                                            const int32_t* __restrict__ frontier,
                                                            const uint32_t* __restrict__ fneed,
                                                            int nf, int32_t* __restrict__ next,
                                                            uint32_t* __restrict__ next_need,
                                                            uint32_t* __restrict__ next_count,
                                                            uint32_t* __restrict__ flag,
                                                            uint32_t* __restrict__ max_need) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nf) return;
    float lo[4][3], hi[4][3];
    int ref[4];
    const int nc = collapse4<true>(nodes, frontier[k], lo, hi, ref);
    const uint32_t need = fneed[k] + (uint32_t)(nc - 1);
    atomicMax(max_need, need);
    for (int c = 0; c < nc; ++c) {
        if (ref[c] < 0) continue;
        flag[ref[c]] = 1u;
        const uint32_t pos = atomicAdd(next_count, 1u);
        next[pos] = ref[c];
        next_need[pos] = need;
    }
}

template <bool GREEDY>
__global__ __launch_bounds__(256) void wide_emit_kernel(int n, const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ idx4,
                                                        const float4* __restrict__ nodes,
                                                        float4* __restrict__ nodes4,
                                                        const uint32_t* __restrict__ bounds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !flag[i]) return;
    const int e_min = wide_min_exponent(bounds);
    float lo[4][3], hi[4][3];
    int ref[4];
    const int nc = collapse4<GREEDY>(nodes, i, lo, hi, ref);
    for (int k = 0; k < nc; ++k)
        if (ref[k] >= 0) ref[k] = (int)idx4[ref[k]];
    float ulo[3] = {INFINITY, INFINITY, INFINITY}, uhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < nc; ++k)
        for (int a = 0; a < 3; ++a) {
            ulo[a] = fminf(ulo[a], lo[k][a]);
            uhi[a] = fmaxf(uhi[a], hi[k][a]);
        }
    float origin[3], scale[3];
    uint32_t meta = (uint32_t)nc << 24;
    for (int a = 0; a < 3; ++a) {
        const int e = wide_exponent(uhi[a] - ulo[a], e_min);
        scale[a] = ldexpf(1.0f, e);
        origin[a] = ulo[a] - scale[a];
        meta |= (uint32_t)(e + 127) << (8 * a);
    }
    float ql[4][3], qh[4][3];
    for (int k = 0; k < nc; ++k) quantize_child(lo[k], hi[k], origin, scale, ql[k], qh[k]);
    for (int k = nc; k < 4; ++k) empty_child(ql, qh, ref, k);
    write_node4(nodes4 + NODE4_F4 * (int64_t)idx4[i], origin, meta, ql, qh, ref);
}

// ------------------------------------------------------------------ host
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

static int dev_alloc(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(WF_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    return WF_OK;
}

// Binned SAH build over the m >= 2 Morton-sorted references: writes child,
// range and parent like karras_kernel and leaves the final reference order in
// `perm` (in place). The large-node levels are driven from the host, which
// reads each level's node list back to cut it into chunks.
static int build_sah(int64_t m64, const float4* refs, uint32_t* perm, int2* child, int2* range,
                     int32_t* parent, hipStream_t s) {
    const int m = (int)m64;
    const int64_t nl_cap = m64 / (SAH_SMALL + 1) + 2;
    const int64_t chunk_cap = m64 / SAH_CHUNK + nl_cap + 2;
    DevBuf large[2], bins, small, chunks, chunk0, lefts, offs, tmp, counters;
    for (int k = 0; k < 2; ++k) WF_TRY(dev_alloc(large[k], sizeof(SahNode) * nl_cap));
    WF_TRY(dev_alloc(bins, sizeof(uint32_t) * SAH_BIN_WORDS * nl_cap));
    WF_TRY(dev_alloc(small, sizeof(int4) * (m64 / 2 + 1)));
    WF_TRY(dev_alloc(chunks, sizeof(int4) * chunk_cap));
    WF_TRY(dev_alloc(chunk0, sizeof(int) * (nl_cap + 1)));
    WF_TRY(dev_alloc(lefts, sizeof(uint32_t) * chunk_cap));
    WF_TRY(dev_alloc(offs, sizeof(int2) * chunk_cap));
    WF_TRY(dev_alloc(tmp, sizeof(uint32_t) * m64));
    WF_TRY(dev_alloc(counters, sizeof(uint32_t) * 4));
    WF_TRY(fill_u32((uint32_t*)counters.p, 4, 0u, s));
    int nl = 0;
    if (m <= SAH_SMALL) {
        const int4 root = make_int4(0, m, -1, 0);
        const uint32_t one = 1;
        WF_HIP(hipMemcpyAsync(small.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
        WF_HIP(hipMemcpyAsync((uint32_t*)counters.p + 1, &one, sizeof(one), hipMemcpyHostToDevice, s));
        WF_HIP(hipStreamSynchronize(s));
    } else {
        SahNode root{};
        root.count = m;
        root.parent = -1;
        for (int a = 0; a < 3; ++a) root.cmin[a] = 0xffffffffu;
        WF_HIP(hipMemcpyAsync(large[0].p, &root, sizeof(root), hipMemcpyHostToDevice, s));
        WF_HIP(hipStreamSynchronize(s));
        nl = 1;
    }
    std::vector<SahNode> hn;
    std::vector<int4> hc;
    std::vector<int> hc0;
    int cur = 0;
    while (nl > 0) {
        if (nl > nl_cap) return fail(WF_ERR_HIP, "wf_scene_create: SAH level larger than its bound");
        hn.resize(nl);
        WF_HIP(hipMemcpyAsync(hn.data(), large[cur].p, sizeof(SahNode) * nl, hipMemcpyDeviceToHost, s));
        WF_HIP(hipStreamSynchronize(s));
        hc.clear();
        hc0.resize(nl + 1);
        for (int j = 0; j < nl; ++j) {
            hc0[j] = (int)hc.size();
            const int end = hn[j].first + hn[j].count;
            for (int st = hn[j].first; st < end; st += SAH_CHUNK)
                hc.push_back(make_int4(j, st, std::min(SAH_CHUNK, end - st), 0));
        }
        hc0[nl] = (int)hc.size();
        const int nc = (int)hc.size();
        if (nc > chunk_cap) return fail(WF_ERR_HIP, "wf_scene_create: SAH chunks beyond their bound");
        WF_HIP(hipMemcpyAsync(chunks.p, hc.data(), sizeof(int4) * nc, hipMemcpyHostToDevice, s));
        WF_HIP(hipMemcpyAsync(chunk0.p, hc0.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice, s));
        const int64_t words = (int64_t)nl * SAH_BIN_WORDS;
        sah_bins_init_kernel<<<grid_for(words, 256), 256, 0, s>>>((uint32_t*)bins.p, words);
        WF_TRY(launch_check("sah_bins_init_kernel"));
        WF_TRY(fill_u32((uint32_t*)counters.p, 1, 0u, s));
        SahNode* nodes = (SahNode*)large[cur].p;
        const int4* ch = (const int4*)chunks.p;
        sah_cbounds_kernel<<<nc, 256, 0, s>>>(ch, perm, refs, nodes);
        WF_TRY(launch_check("sah_cbounds_kernel"));
        sah_bin_kernel<<<nc, 256, 0, s>>>(ch, perm, refs, nodes, (uint32_t*)bins.p);
        WF_TRY(launch_check("sah_bin_kernel"));
        sah_split_kernel<<<grid_for(nl, 4), 256, 0, s>>>(nodes, nl, (const uint32_t*)bins.p, m,
                                                        (SahNode*)large[cur ^ 1].p, (int4*)small.p,
                                                        (uint32_t*)counters.p, child, range, parent);
        WF_TRY(launch_check("sah_split_kernel"));
        sah_count_kernel<<<nc, 256, 0, s>>>(ch, perm, refs, nodes, (uint32_t*)lefts.p);
        WF_TRY(launch_check("sah_count_kernel"));
        sah_offsets_kernel<<<grid_for(nl, 256), 256, 0, s>>>((const int*)chunk0.p, nl,
                                                            (const uint32_t*)lefts.p, (int2*)offs.p);
        WF_TRY(launch_check("sah_offsets_kernel"));
        sah_scatter_kernel<<<nc, 256, 0, s>>>(ch, perm, refs, nodes, (const int2*)offs.p,
                                              (uint32_t*)tmp.p);
        WF_TRY(launch_check("sah_scatter_kernel"));
        sah_copy_kernel<<<nc, 256, 0, s>>>(ch, (const uint32_t*)tmp.p, perm);
        WF_TRY(launch_check("sah_copy_kernel"));
        uint32_t hnext = 0;
        WF_HIP(hipMemcpyAsync(&hnext, counters.p, sizeof(hnext), hipMemcpyDeviceToHost, s));
        WF_HIP(hipStreamSynchronize(s));
        nl = (int)hnext;
        cur ^= 1;
    }
    uint32_t ns = 0;
    WF_HIP(hipMemcpyAsync(&ns, (uint32_t*)counters.p + 1, sizeof(ns), hipMemcpyDeviceToHost, s));
    WF_HIP(hipStreamSynchronize(s));
    if (ns > 0) {
        sah_small_kernel<<<grid_for(ns, SAH_WAVES), 64 * SAH_WAVES, 0, s>>>(
            (const int4*)small.p, (int)ns, refs, perm, m, (const uint32_t*)counters.p, child, range,
            parent);
        WF_TRY(launch_check("sah_small_kernel"));
    }
    WF_HIP(hipStreamSynchronize(s));  // the temporaries leave scope
    return WF_OK;
}

static void free_scene(wf_scene_s* s) {
    if (!s) return;
    if (s->tris) (void)hipFree(s->tris);
    if (s->nodes) (void)hipFree(s->nodes);
    if (s->nodes4) (void)hipFree(s->nodes4);
    if (s->shade) (void)hipFree(s->shade);
    if (s->geom_first) (void)hipFree(s->geom_first);
    if (s->albedo) (void)hipFree(s->albedo);
    if (s->emission) (void)hipFree(s->emission);
    delete s;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_scene_destroy(wf_scene scene) {
    free_scene(scene);
    return WF_OK;
}

extern "C" int wf_scene_set_materials(wf_scene scene, const float* albedo_rgb,
                                      const float* emission_rgb, int num_materials,
                                      wf_stream_t stream) {
    WF_CHECK_ARG(scene, "wf_scene_set_materials: null scene");
    WF_CHECK_ARG(num_materials >= 1 && albedo_rgb && emission_rgb,
                 "wf_scene_set_materials: need >= 1 material and albedo/emission arrays");
    hipStream_t s = as_stream(stream);
    const size_t bytes = sizeof(float) * 3 * (size_t)num_materials;
    if (num_materials != scene->num_materials || !scene->albedo || !scene->emission) {
        float *a = nullptr, *e = nullptr;
        WF_HIP(hipMalloc((void**)&a, bytes));
        if (hipMalloc((void**)&e, bytes) != hipSuccess) {
            (void)hipFree(a);
            return fail(WF_ERR_OUT_OF_MEMORY, "wf_scene_set_materials: out of device memory");
        }
        WF_HIP(hipStreamSynchronize(s));
        if (scene->albedo) (void)hipFree(scene->albedo);
        if (scene->emission) (void)hipFree(scene->emission);
        scene->albedo = a;
        scene->emission = e;
        scene->num_materials = num_materials;
    }
    // host or device source arrays (unified addressing)
    WF_HIP(hipMemcpyAsync(scene->albedo, albedo_rgb, bytes, hipMemcpyDefault, s));
    WF_HIP(hipMemcpyAsync(scene->emission, emission_rgb, bytes, hipMemcpyDefault, s));
    WF_HIP(hipStreamSynchronize(s));
    return WF_OK;
}

extern "C" int wf_scene_get_materials(wf_scene scene, int* num_materials, float* albedo_rgb,
                                      float* emission_rgb) {
    WF_CHECK_ARG(scene, "wf_scene_get_materials: null scene");
    if (num_materials) *num_materials = scene->num_materials;
    const size_t bytes = sizeof(float) * 3 * (size_t)scene->num_materials;
    if (albedo_rgb) WF_HIP(hipMemcpy(albedo_rgb, scene->albedo, bytes, hipMemcpyDefault));
    if (emission_rgb) WF_HIP(hipMemcpy(emission_rgb, scene->emission, bytes, hipMemcpyDefault));
    return WF_OK;
}

extern "C" int wf_scene_create(const void* vertices, int64_t num_vertices, int vertex_stride,
                               const uint32_t* indices, int64_t num_tris,
                               const int32_t* geom_first_tri, int num_geoms,
                               const int32_t* tri_material, wf_stream_t stream, wf_scene* out) {
    WF_CHECK_ARG(out, "wf_scene_create: null out");
    *out = nullptr;
    WF_CHECK_ARG(num_tris >= 0 && num_vertices >= 0, "wf_scene_create: negative size");
    WF_CHECK_ARG(vertex_stride >= 24 && vertex_stride % 4 == 0,
                 "wf_scene_create: vertex_stride must be >= 24 and a multiple of 4");
    WF_CHECK_ARG(num_tris <= (int64_t)(1 << 28), "wf_scene_create: more than 2^28 triangles");
    WF_CHECK_ARG(num_geoms >= 1 && geom_first_tri, "wf_scene_create: need >= 1 geometry");
    WF_CHECK_ARG(geom_first_tri[0] == 0 && geom_first_tri[num_geoms] == num_tris,
                 "wf_scene_create: geom_first_tri must start at 0 and end at num_tris");
    for (int g = 0; g < num_geoms; ++g)
        WF_CHECK_ARG(geom_first_tri[g] <= geom_first_tri[g + 1],
                     "wf_scene_create: geom_first_tri must be ascending");
    if (num_tris > 0) WF_CHECK_ARG(vertices && indices, "wf_scene_create: null vertex/index buffer");

    hipStream_t s = as_stream(stream);
    wf_scene_s* sc = new wf_scene_s();
    int dev = 0;
    (void)hipGetDevice(&dev);
    sc->device = dev;
    sc->num_tris = num_tris;
    sc->num_geoms = num_geoms;
    struct Guard {
        wf_scene_s*& p;
        ~Guard() {
            if (p) free_scene(p);
        }
    };
    wf_scene_s* owned = sc;
    Guard guard{owned};

    const int64_t n = num_tris;
    WF_HIP(hipMalloc((void**)&sc->geom_first, sizeof(int32_t) * (num_geoms + 1)));
    WF_HIP(hipMemcpyAsync(sc->geom_first, geom_first_tri, sizeof(int32_t) * (num_geoms + 1),
                          hipMemcpyHostToDevice, s));
    {
        float a[3] = {0.8f, 0.8f, 0.8f}, e[3] = {0.f, 0.f, 0.f};
        WF_TRY(wf_scene_set_materials(sc, a, e, 1, stream));
    }
    WF_HIP(hipMalloc((void**)&sc->shade, sizeof(float4) * 3 * (n > 0 ? n : 1)));

    if (n == 0) {
        WF_HIP(hipMalloc((void**)&sc->tris, sizeof(float4) * 3));
        WF_HIP(hipMalloc((void**)&sc->nodes, sizeof(float4) * 4));
        WF_HIP(hipMalloc((void**)&sc->nodes4, sizeof(float4) * NODE4_F4));
        // empty nodes: inverted boxes / no children, so any traversal ends at the root
        const float inf = INFINITY;
        const float empty2[16] = {inf, inf, inf, -inf, -inf, -inf, inf, inf,
                                  inf, -inf, -inf, -inf, 0.f, 0.f, 0.f, 0.f};
        // 4-wide root: every child slot empty (planes QMAX above 0, fp16
        // 2047 = 0x67ff), every reference leaf slot 0
        uint32_t empty4[4 * NODE4_F4] = {0u};
        for (int a = 0; a < 3; ++a) empty4[4 + 4 * a] = empty4[5 + 4 * a] = 0x67ff67ffu;
        for (int k = 0; k < 4; ++k) empty4[16 + k] = (uint32_t)leaf_ref(0, 1);
        WF_HIP(hipMemcpy(sc->nodes, empty2, sizeof(empty2), hipMemcpyHostToDevice));
        WF_HIP(hipMemcpy(sc->nodes4, empty4, sizeof(empty4), hipMemcpyHostToDevice));
        sc->num_nodes = 1;
        sc->num_nodes4 = 1;
        sc->num_slots = 0;
        for (int a = 0; a < 3; ++a) sc->bmin[a] = sc->bmax[a] = 0.f;
        sc->num_leaves = 0;
        *out = sc;
        owned = nullptr;
        return WF_OK;
    }

    const Tuning tu = tuning();
    DevBuf tri_box, tri_geom, bounds, misc, counts, offsets, item_box, item_tri;
    WF_TRY(dev_alloc(tri_box, sizeof(float4) * 2 * n));
    WF_TRY(dev_alloc(tri_geom, sizeof(int32_t) * n));
    WF_TRY(dev_alloc(bounds, sizeof(uint32_t) * 16));
    WF_TRY(dev_alloc(misc, sizeof(uint32_t) * 4));  // [0] bad index, [1] depth, [2] leaves

    uint32_t* b = (uint32_t*)bounds.p;
    // min slots start at key(+inf) > everything, max slots at key(-inf) = small
    {
        std::vector<uint32_t> init(16, 0u);
        auto key_host = [](float f) {
            uint32_t u;
            memcpy(&u, &f, 4);
            return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        };
        for (int a = 0; a < 3; ++a) {
            init[a] = key_host(INFINITY);
            init[3 + a] = key_host(-INFINITY);
            init[6 + a] = key_host(INFINITY);
            init[9 + a] = key_host(-INFINITY);
        }
        WF_HIP(hipMemcpyAsync(b, init.data(), sizeof(uint32_t) * 16, hipMemcpyHostToDevice, s));
        WF_HIP(hipStreamSynchronize(s));
    }
    WF_TRY(fill_u32((uint32_t*)misc.p, 4, 0u, s));

    tri_prep_kernel<<<grid_for(n, 256), 256, 0, s>>>(
        (const char*)vertices, num_vertices, vertex_stride, indices, n, sc->geom_first, num_geoms,
        tri_material, (float4*)tri_box.p, sc->shade, (int32_t*)tri_geom.p, b, (uint32_t*)misc.p);
    WF_TRY(launch_check("tri_prep_kernel"));

    // References: one per triangle, or pieces of the large ones.
    int64_t m = n;
    const float4* refs = (const float4*)tri_box.p;
    const uint32_t* ref_tri = nullptr;
    if (tu.bvh_presplit && tu.bvh_split_depth > 0) {
        const float frac = 1.0f / (float)tu.bvh_split_denom;
        WF_TRY(dev_alloc(counts, sizeof(uint32_t) * n));
        split_kernel<false><<<grid_for(n, 128), 128, 0, s>>>(
            (const char*)vertices, vertex_stride, indices, (const float4*)tri_box.p, n, b, frac,
            tu.bvh_split_depth, (uint32_t*)counts.p, nullptr, nullptr, nullptr);
        WF_TRY(launch_check("split_kernel<count>"));
        std::vector<uint32_t> hc(n);
        WF_HIP(hipMemcpyAsync(hc.data(), counts.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
        WF_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> off(n);
        int64_t total = 0;
        for (int64_t t = 0; t < n; ++t) {
            off[t] = total;
            total += hc[t];
        }
        if (total > (int64_t)(1 << 28))
            return fail(WF_ERR_TOO_LARGE, "wf_scene_create: more than 2^28 references after splitting");
        if (total != n) {
            m = total;
            WF_TRY(dev_alloc(offsets, sizeof(int64_t) * n));
            WF_TRY(dev_alloc(item_box, sizeof(float4) * 2 * m));
            WF_TRY(dev_alloc(item_tri, sizeof(uint32_t) * m));
            WF_HIP(hipMemcpyAsync(offsets.p, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
            split_kernel<true><<<grid_for(n, 128), 128, 0, s>>>(
                (const char*)vertices, vertex_stride, indices, (const float4*)tri_box.p, n, b, frac,
                tu.bvh_split_depth, nullptr, (const int64_t*)offsets.p, (float4*)item_box.p,
                (uint32_t*)item_tri.p);
            WF_TRY(launch_check("split_kernel<emit>"));
            WF_HIP(hipStreamSynchronize(s));  // `off` leaves scope
            refs = (const float4*)item_box.p;
            ref_tri = (const uint32_t*)item_tri.p;
        }
    }
    sc->num_slots = m;
    sc->num_nodes = m >= 2 ? m - 1 : 1;
    WF_HIP(hipMalloc((void**)&sc->tris, sizeof(float4) * 3 * m));
    WF_HIP(hipMalloc((void**)&sc->nodes, sizeof(float4) * 4 * sc->num_nodes));

    DevBuf keys, vals, scratch, child, range, parent, box, flags;
    WF_TRY(dev_alloc(keys, sizeof(uint64_t) * m));
    WF_TRY(dev_alloc(vals, sizeof(uint32_t) * m));
    int64_t sb = wf_sort_scratch_bytes(m, 8, 63, 8);
    if (sb < 0) return (int)sb;
    WF_TRY(dev_alloc(scratch, (size_t)sb));
    WF_TRY(dev_alloc(child, sizeof(int2) * (m > 1 ? m - 1 : 1)));
    WF_TRY(dev_alloc(range, sizeof(int2) * (m > 1 ? m - 1 : 1)));
    WF_TRY(dev_alloc(parent, sizeof(int32_t) * (2 * m - 1)));
    WF_TRY(dev_alloc(box, sizeof(float4) * 2 * (2 * m - 1)));
    WF_TRY(dev_alloc(flags, sizeof(uint32_t) * (m > 1 ? m - 1 : 1)));
    if (m > 1) WF_TRY(fill_u32((uint32_t*)flags.p, m - 1, 0u, s));

    if (m == 1) {
        WF_TRY(fill_u32((uint32_t*)vals.p, 1, 0u, s));
        emit_single_kernel<<<1, 1, 0, s>>>(refs, sc->nodes);
        WF_TRY(launch_check("emit_single_kernel"));
        sc->num_leaves = 1;
    } else {
        centroid_keys_kernel<<<grid_for(m, 256), 256, 0, s>>>(refs, m, b, (uint64_t*)keys.p,
                                                              (uint32_t*)vals.p);
        WF_TRY(launch_check("centroid_keys_kernel"));
        WF_TRY(wf_sort_pairs(keys.p, (uint32_t*)vals.p, m, 8, 63, 8, scratch.p, sb, stream, nullptr));
        if (tu.bvh_builder == 1) {
            WF_TRY(build_sah(m, refs, (uint32_t*)vals.p, (int2*)child.p, (int2*)range.p,
                             (int32_t*)parent.p, s));
        } else {
            karras_kernel<<<grid_for(m - 1, 256), 256, 0, s>>>((const uint64_t*)keys.p, (int)m,
                                                               (int2*)child.p, (int2*)range.p,
                                                               (int32_t*)parent.p);
            WF_TRY(launch_check("karras_kernel"));
        }
        refit_kernel<<<grid_for(m, 256), 256, 0, s>>>(
            refs, (const uint32_t*)vals.p, (int)m, (const int2*)child.p, (const int32_t*)parent.p,
            (float4*)box.p, (uint32_t*)flags.p, (uint32_t*)misc.p + 1);
        WF_TRY(launch_check("refit_kernel"));
        emit_nodes_kernel<<<grid_for(m - 1, 256), 256, 0, s>>>(
            (int)m, tu.bvh_leaf_max, (const int2*)child.p, (const int2*)range.p,
            (const float4*)box.p, b, sc->nodes, (uint32_t*)misc.p + 2);
        WF_TRY(launch_check("emit_nodes_kernel"));
    }
    // 4-wide hierarchy
    {
        DevBuf flag4, idx4, scan_state, wide_misc;
        bool greedy = tu.bvh_collapse == 1;
        const int64_t nin = m > 1 ? m - 1 : 1;
        WF_TRY(dev_alloc(wide_misc, sizeof(uint32_t) * 4));  // [0] depth, [1] nodes, [2] greedy stack bound
        WF_TRY(fill_u32((uint32_t*)wide_misc.p, 4, 0u, s));
        uint32_t hw[4] = {0, 1, 0, 0};
        if (m > 1) {
            WF_TRY(dev_alloc(flag4, sizeof(uint32_t) * nin));
            WF_TRY(dev_alloc(idx4, sizeof(uint32_t) * nin));
            const int64_t tiles = (nin + 255) / 256;
            WF_TRY(dev_alloc(scan_state, sizeof(uint32_t) * (1 + tiles)));
            WF_TRY(fill_u32((uint32_t*)scan_state.p, 1 + tiles, 0u, s));
            if (greedy) {
                // top-down: the root, then the interior children of each level
                DevBuf fr[2], fneed[2], fcount;
                for (int k = 0; k < 2; ++k) {
                    WF_TRY(dev_alloc(fr[k], sizeof(int32_t) * nin));
                    WF_TRY(dev_alloc(fneed[k], sizeof(uint32_t) * nin));
                }
                WF_TRY(dev_alloc(fcount, sizeof(uint32_t)));
                WF_TRY(fill_u32((uint32_t*)flag4.p, nin, 0u, s));
                WF_TRY(fill_u32((uint32_t*)flag4.p, 1, 1u, s));
                WF_TRY(fill_u32((uint32_t*)fr[0].p, 1, 0u, s));
                WF_TRY(fill_u32((uint32_t*)fneed[0].p, 1, 0u, s));
                int nf = 1, cur = 0;
                while (nf > 0) {
                    WF_TRY(fill_u32((uint32_t*)fcount.p, 1, 0u, s));
                    wide_frontier_kernel<<<grid_for(nf, 256), 256, 0, s>>>(
                        sc->nodes, (const int32_t*)fr[cur].p, (const uint32_t*)fneed[cur].p, nf,
                        (int32_t*)fr[cur ^ 1].p, (uint32_t*)fneed[cur ^ 1].p, (uint32_t*)fcount.p,
                        (uint32_t*)flag4.p, (uint32_t*)wide_misc.p + 2);
                    WF_TRY(launch_check("wide_frontier_kernel"));
                    uint32_t hn = 0;
                    WF_HIP(hipMemcpyAsync(&hn, fcount.p, sizeof(hn), hipMemcpyDeviceToHost, s));
                    WF_HIP(hipStreamSynchronize(s));
                    nf = (int)hn;
                    cur ^= 1;
                }
                uint32_t need = 0;
                WF_HIP(hipMemcpyAsync(&need, (uint32_t*)wide_misc.p + 2, sizeof(need),
                                      hipMemcpyDeviceToHost, s));
                WF_HIP(hipStreamSynchronize(s));
                // a stack bound beyond the traversal stack: collapse every
                // second level instead
                if (need > (uint32_t)TRACE_STACK4) greedy = false;
            }
            if (!greedy) {
                wide_mark_kernel<<<grid_for(nin, 256), 256, 0, s>>>(
                    (int)m, tu.bvh_leaf_max, (const int2*)range.p, (const int32_t*)parent.p,
                    (uint32_t*)flag4.p, (uint32_t*)wide_misc.p);
                WF_TRY(launch_check("wide_mark_kernel"));
            }
            flag_scan_kernel<<<(unsigned)tiles, 256, 0, s>>>(
                (const uint32_t*)flag4.p, nin, (uint32_t*)idx4.p, (uint32_t*)scan_state.p,
                (uint32_t*)scan_state.p + 1, (uint32_t*)wide_misc.p + 1);
            WF_TRY(launch_check("flag_scan_kernel"));
            WF_HIP(hipMemcpyAsync(hw, wide_misc.p, sizeof(hw), hipMemcpyDeviceToHost, s));
            WF_HIP(hipStreamSynchronize(s));
        }
        sc->num_nodes4 = hw[1];
        // every-second-level collapse: at most 3 pushes per 4-wide level, and
        // levels = real 2-wide depth / 2 + 1; greedy: the largest sum of
        // (children - 1) over a root-to-node path, measured while marking
        sc->max_stack4 = greedy && m > 1 ? (int)hw[2] : 3 * ((int)hw[0] / 2 + 1);
        if (sc->max_stack4 > TRACE_STACK4)
            return fail(WF_ERR_TOO_LARGE, "wf_scene_create: 4-wide BVH deeper than the traversal stack");
        WF_HIP(hipMalloc((void**)&sc->nodes4, sizeof(float4) * NODE4_F4 * sc->num_nodes4));
        if (m > 1) {
            if (greedy)
                wide_emit_kernel<true><<<grid_for(nin, 256), 256, 0, s>>>(
                    (int)m, (const uint32_t*)flag4.p, (const uint32_t*)idx4.p, sc->nodes, sc->nodes4, b);
            else
                wide_emit_kernel<false><<<grid_for(nin, 256), 256, 0, s>>>(
                    (int)m, (const uint32_t*)flag4.p, (const uint32_t*)idx4.p, sc->nodes, sc->nodes4, b);
            WF_TRY(launch_check("wide_emit_kernel"));
        } else {
            wide_single_kernel<<<1, 1, 0, s>>>(sc->nodes, b, sc->nodes4);
            WF_TRY(launch_check("wide_single_kernel"));
        }
        WF_HIP(hipStreamSynchronize(s));  // the temporaries leave scope
    }
    emit_tris_kernel<<<grid_for(m, 256), 256, 0, s>>>(
        (const char*)vertices, vertex_stride, indices, (const uint32_t*)vals.p, ref_tri,
        (const int32_t*)tri_geom.p, sc->geom_first, m, sc->tris);
    WF_TRY(launch_check("emit_tris_kernel"));

    uint32_t hb[16], hm[4];
    WF_HIP(hipMemcpyAsync(hb, b, sizeof(hb), hipMemcpyDeviceToHost, s));
    WF_HIP(hipMemcpyAsync(hm, misc.p, sizeof(hm), hipMemcpyDeviceToHost, s));
    WF_HIP(hipStreamSynchronize(s));
    if (hm[0]) return fail(WF_ERR_INVALID_ARG, "wf_scene_create: index >= num_vertices");
    auto unkey = [](uint32_t k) {
        uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
        float f;
        memcpy(&f, &u, 4);
        return f;
    };
    for (int a = 0; a < 3; ++a) {
        sc->bmin[a] = unkey(hb[a]);
        sc->bmax[a] = unkey(hb[3 + a]);
    }
    sc->max_depth = (int)hm[1];
    if (m > 1) sc->num_leaves = hm[2];
    if (sc->max_depth >= TRACE_STACK)
        return fail(WF_ERR_TOO_LARGE, "wf_scene_create: BVH deeper than the traversal stack");
    *out = sc;
    owned = nullptr;
    return WF_OK;
}

extern "C" int wf_scene_bounds(wf_scene scene, float out_min[3], float out_max[3]) {
    WF_CHECK_ARG(scene && out_min && out_max, "wf_scene_bounds: null argument");
    for (int a = 0; a < 3; ++a) {
        out_min[a] = scene->bmin[a];
        out_max[a] = scene->bmax[a];
    }
    return WF_OK;
}

extern "C" int wf_scene_info(wf_scene scene, int64_t* num_tris, int64_t* num_nodes,
                             int64_t* num_leaves, int* num_geoms) {
    WF_CHECK_ARG(scene, "wf_scene_info: null scene");
    if (num_tris) *num_tris = scene->num_tris;
    if (num_nodes) *num_nodes = scene->num_slots >= 1 ? scene->num_nodes : 0;
    if (num_leaves) *num_leaves = scene->num_leaves;
    if (num_geoms) *num_geoms = scene->num_geoms;
    return WF_OK;
}

extern "C" int wf_scene_wide_info(wf_scene scene, int64_t* num_nodes4, int* max_stack4,
                                  float* nodes4_host) {
    WF_CHECK_ARG(scene, "wf_scene_wide_info: null scene");
    if (num_nodes4) *num_nodes4 = scene->num_nodes4;
    if (max_stack4) *max_stack4 = scene->max_stack4;
    if (nodes4_host)
        WF_HIP(hipMemcpy(nodes4_host, scene->nodes4, sizeof(float4) * NODE4_F4 * scene->num_nodes4,
                         hipMemcpyDeviceToHost));
    return WF_OK;
}

extern "C" int wf_scene_download(wf_scene scene, float* nodes_host, int32_t* tri_order_host) {
    WF_CHECK_ARG(scene, "wf_scene_download: null scene");
    if (scene->num_tris == 0) return WF_OK;
    if (nodes_host)
        WF_HIP(hipMemcpy(nodes_host, scene->nodes, sizeof(float4) * 4 * scene->num_nodes,
                         hipMemcpyDeviceToHost));
    if (tri_order_host) {
        std::vector<float4> t(3 * scene->num_slots);
        WF_HIP(hipMemcpy(t.data(), scene->tris, sizeof(float4) * 3 * scene->num_slots,
                         hipMemcpyDeviceToHost));
        for (int64_t s = 0; s < scene->num_slots; ++s) {
            float w = t[3 * s].w;
            memcpy(&tri_order_host[s], &w, 4);
        }
    }
    return WF_OK;
}
