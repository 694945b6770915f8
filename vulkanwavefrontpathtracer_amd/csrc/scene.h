// Scene = triangles + bounding volume hierarchy + materials, owned by the library.
//
// Device layout (all 16-byte aligned):
//   tris    [3 * num_slots] float4, in BVH leaf order ("slots"; a split
//           large triangle occupies several slots):
//             {v0.xyz, original triangle index}, {v1.xyz, prim in geometry},
//             {v2.xyz, geometry index}   (the .w fields hold int bits)
//   nodes   [max(num_tris - 1, 1)] x 4 float4 (64 bytes):
//             {c0.lo.xyz, c0.hi.x}, {c0.hi.yz, c1.lo.xy}, {c1.lo.z, c1.hi.xyz},
//             {ref0, ref1, 0, 0} (int bits).  ref >= 0: internal node index;
//             ref < 0: leaf ~(first_slot << 3 | (count - 1)), count <= 8.
//   nodes4  [num_nodes4] x NODE4_F4 float4 (80 bytes), the same hierarchy
//           collapsed to 2-4 children per node, the child boxes quantized to
//           integers 0..2047 per plane against the node's own frame and held
//           as fp16 (every such integer is exact in fp16):
//             {origin.xyz, exponents ex | ey << 8 | ez << 16 | nchild << 24},
//             {lo.x c0|c1, lo.x c2|c3, hi.x c0|c1, hi.x c2|c3},
//             the same for y, the same for z   (fp16 pairs, child 0/2 in the
//                                               low half),
//             {ref0, ref1, ref2, ref3};
//           child box plane = origin + q * 2^(e - 127), rounded outwards by
//           at least one quantum (conservative); refs as in nodes, interior
//           refs index nodes4.
//   shade   [3 * num_tris] float4 in ORIGINAL triangle order:
//             {n0.xyz, material}, {n1.xyz, 0}, {n2.xyz, 0}
//   geom_first [num_geoms + 1] int: first original triangle of each geometry.
//   albedo / emission [3 * num_materials] float.
#pragma once

#include "common.h"

// float4 per 4-wide node and the largest plane value of its quantized boxes.
constexpr int NODE4_F4 = 5;
constexpr int NODE4_QMAX = 2047;

struct wf_scene_s {
    int64_t num_tris = 0;
    int64_t num_slots = 0;  // leaf references (>= num_tris with split references)
    int64_t num_nodes = 0;
    int64_t num_leaves = 0;
    int num_geoms = 0;
    int num_materials = 0;
    int max_depth = 0;
    float bmin[3] = {0, 0, 0};
    float bmax[3] = {0, 0, 0};
    float4* tris = nullptr;
    float4* nodes = nullptr;
    float4* nodes4 = nullptr;
    int64_t num_nodes4 = 0;
    int max_stack4 = 0;  // traversal stack bound of nodes4
    float4* shade = nullptr;
    int32_t* geom_first = nullptr;
    float* albedo = nullptr;
    float* emission = nullptr;
    int device = 0;
};

namespace wfpt {

constexpr int TRACE_STACK = 96;        // stack bound of the binary hierarchy's depth checks
constexpr int TRACE_STACK4 = 96;       // traversal stack entries per ray (4-wide)

__host__ __device__ __forceinline__ int leaf_ref(int first, int count) { return ~((first << 3) | (count - 1)); }
__device__ __forceinline__ int leaf_first(int ref) { return (~ref) >> 3; }
__device__ __forceinline__ int leaf_count(int ref) { return ((~ref) & 7) + 1; }

// Material lookup for shading; nmat >= 1.
__device__ __forceinline__ int clamp_material(int m, int nmat) { return (m >= 0 && m < nmat) ? m : 0; }

}  // namespace wfpt
