// Batched rendering: many scenes and views of uneven size in one persistent
// wavefront queue (batch.h / batch.hip).  No counterpart in the reference,
// whose Renderer::render (Renderer.cpp:552-654) renders one view of one
// scene with a host round trip per bounce.
//
// Views are ordered by descending spp so that the views taking pass p form a
// prefix of the global pixel range; a frame is, per pass, one tiny kernel
// that sets the pass index and primary count on the device, then the pass's
// fixed launch sequence, captured once in a HIP graph and replayed.  Sort
// digit width, trace block size and reorder period are chosen by
// wf_batch_autotune, which times one full pass of every candidate.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch.h"
#include "common.h"
#include "host.h"

struct wf_batch_s {
    hipStream_t stream = nullptr;
    int depth = 6, shadow_samples = 2;
    std::vector<wfpt::BatchScene> scenes;
    std::vector<wf_view> views;       // in device order (descending spp)
    std::vector<int> user_to_dev;     // caller's view index -> device order
    std::vector<int64_t> offsets;     // first global pixel per device view (+ total)
    int64_t total_pixels = 0;
    int max_spp = 1;
    std::vector<uint32_t> pass_count;  // primary rays per pass
    wfpt::DeviceBuffer d_scenes, d_views, d_offsets;
    wfpt::DeviceBuffer rays[2], s2i[2], hits, shadow_rays, shadow_hits, keys, idx, sort_scratch,
        compact, seeds, dec, aux, frame, q;
    wfpt::BatchTuning tuning;
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    // What the captured graph was recorded with: the batch tuning, the sort
    // scratch it points at and the process-wide knobs' generation.
    wfpt::BatchTuning graph_tuning;
    void* graph_scratch = nullptr;
    uint64_t graph_generation = 0;
    uint64_t totals[3] = {0, 0, 0};

    ~wf_batch_s() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace wfpt {

static BatchTables tables(const wf_batch_s* b) {
    BatchTables t;
    t.scenes = b->d_scenes.as<BatchScene>();
    t.views = b->d_views.as<BatchView>();
    t.offsets = b->d_offsets.as<int64_t>();
    t.num_views = (int)b->views.size();
    t.shadow_samples = b->shadow_samples;
    return t;
}

static BatchBuffers buffers(wf_batch_s* b) {
    BatchBuffers bb;
    for (int k = 0; k < 2; ++k) {
        bb.rays[k] = b->rays[k].as<wf_ray>();
        bb.s2i[k] = b->s2i[k].as<int32_t>();
    }
    bb.hits = b->hits.as<wf_hit>();
    bb.shadow_rays = b->shadow_rays.as<wf_ray>();
    bb.shadow_hits = b->shadow_hits.as<wf_hit>();
    bb.keys = b->keys.as<uint64_t>();
    bb.idx = b->idx.as<uint32_t>();
    bb.sort_scratch = b->sort_scratch.p;
    bb.sort_scratch_bytes = (int64_t)b->sort_scratch.bytes;
    bb.compact = b->compact.as<uint32_t>();
    bb.seeds = b->seeds.as<uint32_t>();
    bb.dec = b->dec.as<float4>();
    bb.aux = b->aux.as<float4>();
    bb.frame = b->frame.as<float4>();
    bb.q = b->q.as<uint32_t>();
    bb.capacity = b->total_pixels;
    return bb;
}

static int drop_graph(wf_batch_s* b) {
    if (b->graph) {
        WF_HIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    return WF_OK;
}

// Growing the scratch frees the old allocation, which a captured graph may
// still point at: the graph goes with it.
static int ensure_sort_scratch(wf_batch_s* b, int digit_bits) {
    const int64_t need = wf_sort_scratch_bytes(b->total_pixels, 8, 64, digit_bits);
    if (need < 0) return (int)need;
    if ((size_t)need > b->sort_scratch.bytes) WF_TRY(drop_graph(b));
    return b->sort_scratch.ensure((size_t)need);
}

// Capture one pass for the current tuning (replayed for every pass).
static int capture_pass(wf_batch_s* b) {
    WF_TRY(drop_graph(b));
    const BatchTables t = tables(b);
    const BatchBuffers bb = buffers(b);
    hipGraph_t g = nullptr;
    WF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = batch_enqueue_pass(t, bb, b->tuning, b->depth, b->stream);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != WF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(WF_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return fail(WF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    b->graph_tuning = b->tuning;
    b->graph_scratch = b->sort_scratch.p;
    b->graph_generation = tuning_generation();
    return WF_OK;
}

static bool same_tuning(const BatchTuning& a, const BatchTuning& c) {
    return a.digit_bits == c.digit_bits && a.block == c.block &&
           a.reorder_period == c.reorder_period;
}

static int run_pass(wf_batch_s* b, int frame, int pass) {
    const BatchBuffers bb = buffers(b);
    WF_TRY(batch_set_pass(bb, frame, pass, b->pass_count[pass], b->depth, b->stream));
    if (b->use_graph) {
        if (!b->graph || !same_tuning(b->graph_tuning, b->tuning) ||
            b->graph_scratch != b->sort_scratch.p || b->graph_generation != tuning_generation())
            WF_TRY(capture_pass(b));
        WF_HIP(hipGraphLaunch(b->graph, b->stream));
        return WF_OK;
    }
    return batch_enqueue_pass(tables(b), bb, b->tuning, b->depth, b->stream);
}

static int render_batch(wf_batch_s* b, float* ms) {
    WF_TRY(ensure_sort_scratch(b, b->tuning.digit_bits));
    StageTimer t(b->stream, ms);
    WF_TRY(batch_init_frame(tables(b), buffers(b), b->total_pixels, 1, b->stream));
    for (int p = 0; p < b->max_spp; ++p) WF_TRY(run_pass(b, 1, p));
    WF_TRY(t.finish());
    uint64_t tot[3];
    WF_HIP(hipMemcpyAsync(tot, b->q.as<uint32_t>() + BQ_TOTALS64, sizeof(tot),
                          hipMemcpyDeviceToHost, b->stream));
    WF_HIP(hipStreamSynchronize(b->stream));
    for (int k = 0; k < 3; ++k) b->totals[k] = tot[k];
    return WF_OK;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_batch_create(const wf_scene* scenes, int num_scenes, const wf_view* views,
                               int num_views, int depth, int shadow_samples, wf_batch* out) {
    WF_CHECK_ARG(out, "wf_batch_create: null out");
    *out = nullptr;
    WF_CHECK_ARG(scenes && views, "wf_batch_create: null scenes or views");
    WF_CHECK_ARG(num_scenes >= 1 && num_scenes <= 64,
                 "wf_batch_create: 1..64 scenes (6-bit scene prefix in the sort key)");
    WF_CHECK_ARG(num_views >= 1, "wf_batch_create: no views");
    WF_CHECK_ARG(depth >= 0 && depth <= BQ_MAX_DEPTH, "wf_batch_create: depth out of range");
    WF_CHECK_ARG(shadow_samples >= 1 && shadow_samples <= 64, "wf_batch_create: shadow_samples in [1, 64]");
    int64_t total = 0;
    for (int v = 0; v < num_views; ++v) {
        const wf_view& w = views[v];
        WF_CHECK_ARG(w.scene >= 0 && w.scene < num_scenes && scenes[w.scene],
                     "wf_batch_create: view names a missing scene");
        WF_CHECK_ARG(w.width > 0 && w.height > 0 && w.spp >= 1, "wf_batch_create: empty view");
        WF_CHECK_ARG(w.fov > 0 && w.fov < 180 && w.znear > 0 && w.zfar > w.znear,
                     "wf_batch_create: bad camera");
        total += (int64_t)w.width * w.height;
    }
    if (total * shadow_samples > (int64_t)WF_MAX_ELEMENTS)
        return fail(WF_ERR_TOO_LARGE, "wf_batch_create: pixels x shadow samples exceed 2^30-1");
    auto b = std::make_unique<wf_batch_s>();
    b->depth = depth;
    b->shadow_samples = shadow_samples;
    std::vector<int> order(num_views);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return views[a].spp > views[c].spp; });
    b->user_to_dev.assign(num_views, 0);
    std::vector<BatchView> dv;
    b->offsets.push_back(0);
    for (int k = 0; k < num_views; ++k) {
        const wf_view& w = views[order[k]];
        b->user_to_dev[order[k]] = k;
        b->views.push_back(w);
        BatchView bv{};
        CameraMatrices cm;
        camera_matrices(w.position, w.rotation, w.fov, (float)w.width / (float)w.height, w.znear,
                        w.zfar, cm);
        std::memcpy(bv.s2w.m, cm.screen_to_world, sizeof(bv.s2w.m));
        for (int a = 0; a < 3; ++a) {
            bv.origin[a] = w.position[a];
            bv.light[a] = w.light[a];
            bv.background[a] = w.background[a];
        }
        bv.radius = w.shadow_radius;
        bv.max_dist = w.zfar;
        bv.width = w.width;
        bv.height = w.height;
        bv.spp = w.spp;
        bv.scene = w.scene;
        dv.push_back(bv);
        b->offsets.push_back(b->offsets.back() + (int64_t)w.width * w.height);
    }
    b->total_pixels = b->offsets.back();
    b->max_spp = b->views.front().spp;
    for (int p = 0; p < b->max_spp; ++p) {
        int64_t c = 0;
        for (int k = 0; k < num_views; ++k)
            if (b->views[k].spp > p) c = b->offsets[k + 1];
        b->pass_count.push_back((uint32_t)c);
    }
    for (int s = 0; s < num_scenes; ++s) {
        const wf_scene sc = scenes[s];
        BatchScene bs{};
        bs.nodes = sc->nodes;
        bs.nodes4 = sc->nodes4;
        bs.tris = sc->tris;
        bs.shade = sc->shade;
        bs.geom_first = sc->geom_first;
        bs.albedo = sc->albedo;
        bs.emission = sc->emission;
        bs.nmat = sc->num_materials;
        for (int a = 0; a < 3; ++a) {
            bs.bmin[a] = sc->bmin[a];
            bs.bsize[a] = sc->bmax[a] - sc->bmin[a];
        }
        b->scenes.push_back(bs);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(WF_ERR_HIP, "hipStreamCreate failed");
    const int64_t n = b->total_pixels, S = shadow_samples;
    WF_TRY(b->d_scenes.ensure(sizeof(BatchScene) * b->scenes.size()));
    WF_TRY(b->d_views.ensure(sizeof(BatchView) * dv.size()));
    WF_TRY(b->d_offsets.ensure(sizeof(int64_t) * b->offsets.size()));
    WF_HIP(hipMemcpy(b->d_scenes.p, b->scenes.data(), sizeof(BatchScene) * b->scenes.size(), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(b->d_views.p, dv.data(), sizeof(BatchView) * dv.size(), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(b->d_offsets.p, b->offsets.data(), sizeof(int64_t) * b->offsets.size(), hipMemcpyHostToDevice));
    for (int k = 0; k < 2; ++k) {
        WF_TRY(b->rays[k].ensure(sizeof(wf_ray) * n));
        WF_TRY(b->s2i[k].ensure(sizeof(int32_t) * n));
    }
    WF_TRY(b->hits.ensure(sizeof(wf_hit) * n));
    WF_TRY(b->shadow_rays.ensure(sizeof(wf_ray) * n * S));
    WF_TRY(b->shadow_hits.ensure(sizeof(wf_hit) * n * S));
    WF_TRY(b->keys.ensure(sizeof(uint64_t) * n));
    WF_TRY(b->idx.ensure(sizeof(uint32_t) * n));
    WF_TRY(b->compact.ensure(sizeof(uint32_t) * (64 + (n + 255) / 256)));
    WF_TRY(b->seeds.ensure(sizeof(uint32_t) * n));
    WF_TRY(b->dec.ensure(16 * n));
    WF_TRY(b->aux.ensure(16 * n));
    WF_TRY(b->frame.ensure(16 * n));
    WF_TRY(b->q.ensure(sizeof(uint32_t) * BQ_WORDS));
    WF_TRY(ensure_sort_scratch(b.get(), b->tuning.digit_bits));
    *out = b.release();
    return WF_OK;
}

extern "C" int wf_batch_destroy(wf_batch b) {
    delete b;
    return WF_OK;
}

extern "C" int wf_batch_set_tuning(wf_batch b, int digit_bits, int block, int reorder_period) {
    WF_CHECK_ARG(b, "wf_batch_set_tuning: null batch");
    WF_CHECK_ARG(digit_bits == 4 || digit_bits == 8 || digit_bits == 11,
                 "wf_batch_set_tuning: digit_bits must be 4, 8 or 11");
    WF_CHECK_ARG(block == 64 || block == 128 || block == 256,
                 "wf_batch_set_tuning: block must be 64, 128 or 256");
    WF_CHECK_ARG(reorder_period >= 0, "wf_batch_set_tuning: reorder_period < 0 (0 = never)");
    b->tuning.digit_bits = digit_bits;
    b->tuning.block = block;
    b->tuning.reorder_period = reorder_period;
    return ensure_sort_scratch(b, digit_bits);
}

extern "C" int wf_batch_get_tuning(wf_batch b, int* digit_bits, int* block, int* reorder_period) {
    WF_CHECK_ARG(b, "wf_batch_get_tuning: null batch");
    if (digit_bits) *digit_bits = b->tuning.digit_bits;
    if (block) *block = b->tuning.block;
    if (reorder_period) *reorder_period = b->tuning.reorder_period;
    return WF_OK;
}

extern "C" int wf_batch_set_graph(wf_batch b, int enable) {
    WF_CHECK_ARG(b, "wf_batch_set_graph: null batch");
    b->use_graph = enable != 0;
    return WF_OK;
}

extern "C" int wf_batch_render(wf_batch b, float* ms) {
    WF_CHECK_ARG(b, "wf_batch_render: null batch");
    return render_batch(b, ms);
}

extern "C" int wf_batch_autotune(wf_batch b, float* best_ms) {
    WF_CHECK_ARG(b, "wf_batch_autotune: null batch");
    const int digits[3] = {4, 8, 11}, blocks[3] = {64, 128, 256}, periods[3] = {1, 2, 4};
    BatchTuning best = b->tuning;
    float best_t = 3.0e38f;
    WF_TRY(batch_init_frame(tables(b), buffers(b), b->total_pixels, 1, b->stream));
    for (int d : digits)
        for (int k : blocks)
            for (int p : periods) {
                b->tuning.digit_bits = d;
                b->tuning.block = k;
                b->tuning.reorder_period = p;
                WF_TRY(ensure_sort_scratch(b, d));
                // pass 0 is the largest; run it twice and time the second
                WF_TRY(run_pass(b, 1, 0));
                float ms = 0.f;
                {
                    StageTimer tt(b->stream, &ms);
                    WF_TRY(run_pass(b, 1, 0));
                    WF_TRY(tt.finish());
                }
                if (ms < best_t) {
                    best_t = ms;
                    best = b->tuning;
                }
            }
    b->tuning = best;
    WF_TRY(ensure_sort_scratch(b, best.digit_bits));
    WF_TRY(drop_graph(b));
    if (best_ms) *best_ms = best_t;
    return WF_OK;
}

extern "C" int wf_batch_read_pixels(wf_batch b, int view, float* host_rgba) {
    WF_CHECK_ARG(b && host_rgba, "wf_batch_read_pixels: null argument");
    WF_CHECK_ARG(view >= 0 && view < (int)b->views.size(), "wf_batch_read_pixels: bad view");
    const int k = b->user_to_dev[view];
    const int64_t n = b->offsets[k + 1] - b->offsets[k];
    WF_HIP(hipStreamSynchronize(b->stream));
    WF_HIP(hipMemcpy(host_rgba, b->frame.as<float4>() + b->offsets[k], 16 * n, hipMemcpyDeviceToHost));
    return WF_OK;
}

extern "C" int wf_batch_stats(wf_batch b, uint64_t counts[3], int64_t* total_pixels) {
    WF_CHECK_ARG(b, "wf_batch_stats: null batch");
    if (counts)
        for (int k = 0; k < 3; ++k) counts[k] = b->totals[k];
    if (total_pixels) *total_pixels = b->total_pixels;
    return WF_OK;
}
