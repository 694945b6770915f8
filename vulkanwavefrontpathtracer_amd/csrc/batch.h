// Batched wavefront queue: many scenes and views of uneven size in one ray
// queue.  Device-side tables describe the scenes and views; every ray finds
// its view from its global pixel index (views occupy consecutive pixel
// ranges) and its scene from the view.  Counts stay on the device, so one
// pass of the bounce loop is a fixed launch sequence that can be captured
// in a HIP graph and replayed.
#pragma once

#include "scene.h"
#include "shading.h"

namespace wfpt {

struct BatchScene {
    const float4* nodes;
    const float4* nodes4;
    const float4* tris;
    const float4* shade;
    const int32_t* geom_first;
    const float* albedo;
    const float* emission;
    int nmat;
    float bmin[3], bsize[3];  // key frame (scene AABB)
};

struct BatchView {
    Mat4 s2w;
    float origin[3], light[3], background[3];
    float radius, max_dist;
    int width, height, spp, scene;
};

// Pointers and sizes shared by every batched kernel of one queue.
struct BatchTables {
    const BatchScene* scenes;
    const BatchView* views;
    const int64_t* offsets;  // [num_views + 1] first global pixel of each view
    int num_views;
    int shadow_samples;
};

// Device words of the queue: pass index, frame index, counts[b] = rays at
// vertex b of the current pass (b = 0: primary rays), and the frame's
// running ray totals as 64-bit counters.
enum {
    BQ_PASS = 0,
    BQ_FRAME = 1,
    BQ_COUNT0 = 8,      // counts[b] at word BQ_COUNT0 + b, b <= BQ_MAX_DEPTH + 1
    BQ_MAX_DEPTH = 30,
    BQ_TOTALS64 = 48,   // u64 totals {primary, shadow, path} from word 48
    BQ_WORDS = 64
};

struct BatchBuffers {
    wf_ray* rays[2];
    int32_t* s2i[2];
    wf_hit* hits;
    wf_ray* shadow_rays;
    wf_hit* shadow_hits;
    uint64_t* keys;
    uint32_t* idx;
    void* sort_scratch;
    int64_t sort_scratch_bytes;
    uint32_t* compact;  // tile counter + look-back words
    uint32_t* seeds;
    float4* dec;
    float4* aux;
    float4* frame;
    uint32_t* q;  // BQ_WORDS counters
    int64_t capacity;
};

struct BatchTuning {
    int digit_bits = 8, block = 64, reorder_period = 1;
};

// Enqueue one pass (primary rays of every view with spp > pass, then
// `depth` bounces) on stream s.  q[BQ_PASS] and q[BQ_COUNT0] must hold the
// pass index and its primary ray count.
int batch_enqueue_pass(const BatchTables& t, const BatchBuffers& b, const BatchTuning& tu,
                       int depth, hipStream_t s);
int batch_init_frame(const BatchTables& t, const BatchBuffers& b, int64_t total_pixels, int frame,
                     hipStream_t s);
int batch_set_pass(const BatchBuffers& b, int frame, int pass, uint32_t count, int depth,
                   hipStream_t s);

}  // namespace wfpt
