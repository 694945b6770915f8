// Parameters of the path-replay adjoint launch (adjoint.hip), filled by the
// renderer from its settings and the last rendered frame.
#pragma once

#include "scene.h"
#include "shading.h"

namespace wfpt {

constexpr int ADJ_MAX_VERTS = 8;                 // path vertices with an albedo factor (depth <= 8)
constexpr size_t ADJ_LDS_GRAD_BYTES = 32 * 1024;  // gradients kept in LDS up to 1365 materials

struct AdjointParams {
    const float4* nodes;
    const float4* nodes4;
    const float4* tris;
    const float4* shade;
    const float* albedo;
    const float* emission;
    int nmat;
    Mat4 s2w;
    float origin[3];
    const int32_t* index_to_pixel;
    int width, height, spp, depth, frame, S;
    float light[3], radius, background[3], max_dist;
    const float4* d_pixels;   // dLoss/dpixel (rgb used)
    float* grad_albedo;       // [3 * nmat], accumulated
    float* grad_emission;     // [3 * nmat], accumulated
    float4* replay;           // optional: the replayed image
};

int adjoint_launch(const AdjointParams& p, hipStream_t s);

}  // namespace wfpt
