// Frame loop: RayBuffer, PathQueue and Renderer of the reference
// (projects/WaveFrontPathTracer/Ray/RayBuffer.cpp, Renderer/Renderer.cpp)
// driving the HIP stages through the C ABI on one HIP stream.
//
// Differences from the reference, all deliberate:
//   - stage times are hipEvent pairs recorded without waiting; the frame
//     waits once at its end (the reference waits on a fence per dispatch);
//   - the shadow-hit statistics are accumulated on the device and read once
//     per frame;
//   - random numbers are keyed by pixel and compaction keeps input order,
//     so sorting and reordering do not change the image.
#include <algorithm>
#include <cstring>

#include "adjoint.h"
#include "common.h"
#include "host.h"
#include "scene.h"

namespace wfpt {

// ------------------------------------------------------------- RayBuffer
class RayBuffer {
  public:
    DeviceBuffer rays[2], i2s[2], s2i[2], results, keys, indices, scratch;
    int64_t size = 0, capacity = 0, index_capacity = 0;
    int swap = 0;
    bool closest = true;

    wf_ray* ray_buf() { return rays[swap].as<wf_ray>(); }
    int32_t* s2i_buf() { return s2i[swap].as<int32_t>(); }
    int32_t* i2s_buf() { return i2s[swap].as<int32_t>(); }
    wf_hit* hit_buf() { return results.as<wf_hit>(); }

    // RayBuffer::resize (RayBuffer.cpp:49-65): grow the active half only.
    int resize(int64_t n, int64_t index_cap) {
        if (n > capacity || index_cap > index_capacity) {
            capacity = std::max(capacity, n);
            index_capacity = std::max({index_capacity, index_cap, capacity});
            WF_TRY(rays[swap].ensure(sizeof(wf_ray) * capacity));
            WF_TRY(results.ensure(sizeof(wf_hit) * capacity));
            WF_TRY(i2s[swap].ensure(sizeof(int32_t) * index_capacity));
            WF_TRY(s2i[swap].ensure(sizeof(int32_t) * capacity));
        }
        size = n;
        return WF_OK;
    }
    // RayBuffer::resizeReorderingBuffers (RayBuffer.cpp:67-82) plus the sort
    // scratch that Tracer::radixSort grows (Tracer.cpp:43-51).
    int prepare_sort(bool reorder, int key_bytes, int key_bits, int digit_bits) {
        int o = 1 - swap;
        if (reorder) {
            WF_TRY(rays[o].ensure(sizeof(wf_ray) * capacity));
            WF_TRY(i2s[o].ensure(sizeof(int32_t) * index_capacity));
            WF_TRY(s2i[o].ensure(sizeof(int32_t) * capacity));
        }
        WF_TRY(keys.ensure((size_t)key_bytes * capacity));
        WF_TRY(indices.ensure(sizeof(uint32_t) * capacity));
        int64_t sb = wf_sort_scratch_bytes(capacity, key_bytes, key_bits, digit_bits);
        if (sb < 0) return (int)sb;
        WF_TRY(scratch.ensure((size_t)sb));
        return WF_OK;
    }
    wf_ray_buffer view() {
        wf_ray_buffer v{};
        for (int h = 0; h < 2; ++h) {
            v.rays[h] = rays[h].as<wf_ray>();
            v.index_to_slot[h] = i2s[h].as<int32_t>();
            v.slot_to_index[h] = s2i[h].as<int32_t>();
        }
        v.results = results.as<wf_hit>();
        v.keys = keys.p;
        v.indices = indices.as<uint32_t>();
        v.sort_scratch = scratch.p;
        v.sort_scratch_bytes = (int64_t)scratch.bytes;
        v.size = size;
        v.capacity = capacity;
        v.swap = swap;
        v.closest_hit = closest ? 1 : 0;
        return v;
    }
};

// ---------------------------------------------------------------- timing
enum Stage {
    ST_RAYGEN = 0,
    ST_KEYS,
    ST_SORT,
    ST_REORDER,
    ST_TRACE,
    ST_TRACE_UNSORTED,
    ST_SHADE,
    ST_RECONSTRUCT,
    ST_COUNT
};
enum RayKind { RK_PRIMARY = 0, RK_SHADOW, RK_PATH };

class EventLog {
  public:
    struct Rec {
        hipEvent_t a, b;
        int stage, kind;
        int bounce;   // path bounce (-1 for the primary vertex)
        int64_t n;    // rays the stage processed
        bool sorted;  // a trace of a sorted wavefront
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(hipStream_t s, int stage, int kind, int bounce = -1, int64_t n = 0,
               bool sorted = false) {
        Rec r{get(), get(), stage, kind, bounce, n, sorted};
        (void)hipEventRecord(r.a, s);
        recs.push_back(r);
    }
    void end(hipStream_t s) { (void)hipEventRecord(recs.back().b, s); }
    void reset() {
        recs.clear();
        used = 0;
    }
    ~EventLog() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace wfpt

// ----------------------------------------------------------------- renderer
struct wf_renderer_s {
    wfpt::Environment env;
    hipStream_t stream = nullptr;
    wf_scene scene = nullptr;
    bool own_scene = false;
    float smin[3] = {0, 0, 0}, smax[3] = {1, 1, 1};
    // settings (Renderer.cpp:300-384 and the env)
    int width = 1024, height = 768;
    std::string ray_type = "primary";
    float key_value = 0.6f, white_point = 2.0f, shadow_radius = 0.001f;
    int spp = 1, shadow_samples = 2, depth = 1;
    bool sort_shadow = false, reorder_shadow = false, sort_path = false, reorder_path = false;
    bool print_sort_logs = false, headlight = false, russian_roulette = false;
    int rr_start = 3, key_mode = WF_KEY_AILA32, digit_bits = 8, reorder_period = 1;
    float light[3] = {5, -2, 1}, background[3] = {0, 0, 1};
    float cam_pos[3] = {0, 0, 0}, cam_rot[3] = {0, 0, 0}, znear = 0.001f, zfar = 3.f, fov = 45.f;
    int frame_index = 1;
    // prefix of a relative Scene.filename (getAssetPath, base/VulkanTools.cpp:13-29)
    std::string asset_path = "./../assets/";
    // buffers
    wfpt::DeviceBuffer pixels, frame_pixels, aux_pixels, decreases, seeds, i2p, counters, compact;
    wfpt::RayBuffer primary, shadow, queue[2];
    int queue_swap = 0;
    int lut_w = 0, lut_h = 0;
    wfpt::EventLog log;
    // stats of the last frame
    uint64_t n_rays[3] = {0, 0, 0};
    uint64_t sorted_rays = 0;
    float trace_ms[3] = {0, 0, 0};
    float stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // per-bounce sort logs of the last frame (SortLog, Renderer.h:29-46):
    // path bounces 0..depth-1, shadow vertices 0..depth
    std::vector<wf_sort_log> path_logs, shadow_logs;
    // decision log: per pass and path vertex, the triangle hit by each
    // pixel's ray (-2 = no ray) and per shadow sample the light's visibility
    bool decision_log = false;
    wfpt::DeviceBuffer dlog_tri, dlog_vis, dlog_sign;
    // capture of pre-sort path wavefronts
    bool capture = false;
    std::vector<std::unique_ptr<wfpt::DeviceBuffer>> cap;
    std::vector<int64_t> cap_n;

    ~wf_renderer_s() {
        if (own_scene && scene) wf_scene_destroy(scene);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace wfpt {

// The stream is created on first GPU use so that configuration can be read
// and validated on a host without a device.
static int ensure_stream(wf_renderer_s* r) {
    if (r->stream) return WF_OK;
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(WF_ERR_HIP, "hipStreamCreate failed");
    return WF_OK;
}

static int load_settings(wf_renderer_s* r) {
    Environment& e = r->env;
    e.get_int("Resolution.width", r->width);
    e.get_int("Resolution.height", r->height);
    e.get_string("Renderer.rayType", r->ray_type);
    // Renderer setters (Renderer.cpp:390-476) keep the previous value for
    // out-of-range input.
    float f;
    int i;
    if (e.get_float("Renderer.keyValue", f) && f > 0 && f <= 2.0f) r->key_value = f;
    if (e.get_float("Renderer.whitePoint", f) && f > 0 && f <= 2.0f) r->white_point = f;
    if (e.get_float("Renderer.shadowRadius", f) && f > 0 && f <= 100.0f) r->shadow_radius = f;
    if (e.get_int("Renderer.numberOfPrimarySamples", i) && i > 0 && i <= 1024) r->spp = i;
    if (e.get_int("Renderer.numberOfShadowSamples", i) && i > 0 && i <= 1024) r->shadow_samples = i;
    if (e.get_int("Renderer.recursionDepth", i) && i >= 0 && i <= 8) r->depth = i;
    e.get_bool("Renderer.sortShadowRays", r->sort_shadow);
    e.get_bool("Renderer.reorderShadowRays", r->reorder_shadow);
    e.get_bool("Renderer.sortPathRays", r->sort_path);
    e.get_bool("Renderer.reorderPathRays", r->reorder_path);
    e.get_bool("Renderer.russianRoulette", r->russian_roulette);
    e.get_bool("Benchmark.printSortLogs", r->print_sort_logs);
    e.get_bool("Scene.headlight", r->headlight);
    e.get_vector("Scene.light", r->light);
    e.get_vector("Scene.backgroundcolor", r->background);
    e.get_vector("Camera.position", r->cam_pos);
    e.get_vector("Camera.rotation", r->cam_rot);
    e.get_float("Camera.nearPlane", r->znear);
    e.get_float("Camera.farPlane", r->zfar);
    e.get_float("Camera.fieldOfView", r->fov);
    if (e.get_int("Renderer.russianRouletteStart", i) && i >= 0) r->rr_start = i;
    if (e.get_int("Renderer.digitBits", i) && (i == 4 || i == 8 || i == 11)) r->digit_bits = i;
    if (e.get_int("Renderer.reorderPeriod", i) && i >= 0) r->reorder_period = i;
    std::string km;
    if (e.get_string("Renderer.keyMode", km)) {
        if (km == "aila32") r->key_mode = WF_KEY_AILA32;
        else if (km == "origin30") r->key_mode = WF_KEY_ORIGIN30;
        else if (km == "origin30_oct30") r->key_mode = WF_KEY_ORIGIN30_OCT30;
        else if (km == "origin_dir_morton60") r->key_mode = WF_KEY_ORIGIN_DIR_MORTON60;
        else if (km == "origin_dir_morton30") r->key_mode = WF_KEY_ORIGIN_DIR_MORTON30;
        else return fail(WF_ERR_INVALID_ARG,
                         "Renderer.keyMode must be aila32, origin30, origin30_oct30, origin_dir_morton60 "
                         "or origin_dir_morton30");
    }
    if (r->width <= 0 || r->height <= 0 || (int64_t)r->width * r->height > (int64_t)WF_MAX_ELEMENTS)
        return fail(WF_ERR_INVALID_ARG, "Resolution out of range");
    if (r->ray_type != "primary" && r->ray_type != "shadow" && r->ray_type != "path")
        r->ray_type = "primary";  // Renderer::stringToRayType default
    if (r->headlight)
        for (int a = 0; a < 3; ++a) r->light[a] = r->cam_pos[a];
    return WF_OK;
}

static int create(const std::string& text, bool is_path, wf_renderer* out) {
    WF_CHECK_ARG(out, "wf_renderer_create: null out");
    *out = nullptr;
    auto* r = new wf_renderer_s();
    int st = is_path ? r->env.read_file(text) : r->env.read_string(text);
    if (st == WF_OK) st = load_settings(r);
    if (is_path) {
        // env/<name>.json beside assets/, as in the reference's tree
        const size_t slash = text.find_last_of('/');
        r->asset_path = (slash == std::string::npos ? std::string(".") : text.substr(0, slash)) +
                        "/../assets/";
    }
    if (st != WF_OK) {
        delete r;
        return st;
    }
    *out = r;
    return WF_OK;
}

// ---------------------------------------------------------------- passes
struct Frame {
    wf_renderer_s* r;
    hipStream_t s;
    wf_stream_t ws;
    int bounce = -1;
    int pass = 0;
    int64_t cur_n = 0;
    bool cur_sorted = false;

    void tbegin(int stage, int kind) { r->log.begin(s, stage, kind, bounce, cur_n, cur_sorted); }

    // decision log slots of (pass, vertex): vertex 0 = primary hit, b + 1 =
    // the hit of path bounce b
    int32_t* dlog_tri(int vertex) {
        const int64_t np = (int64_t)r->width * r->height;
        return r->dlog_tri.as<int32_t>() + ((int64_t)pass * (r->depth + 1) + vertex) * np;
    }
    uint8_t* dlog_sign(int vertex) {
        const int64_t np = (int64_t)r->width * r->height;
        return r->dlog_sign.as<uint8_t>() + ((int64_t)pass * (r->depth + 1) + vertex) * np;
    }
    uint8_t* dlog_vis(int vertex) {
        const int64_t np = (int64_t)r->width * r->height;
        return r->dlog_vis.as<uint8_t>() +
               ((int64_t)pass * (r->depth + 1) + vertex) * np * r->shadow_samples;
    }
    void tend() { r->log.end(s); }

    // Tracer::trace / Tracer::traceSort (Tracer.cpp:185-225) on a RayBuffer.
    int trace_rays(RayBuffer& rb, int kind, bool sort, bool reorder) {
        int64_t n = rb.size;
        if (n == 0) return WF_OK;
        cur_n = n;
        cur_sorted = false;
        if (sort && kind == RK_PATH && r->reorder_period != 1) {
            // sort only every reorder_period-th bounce (0 = never)
            sort = r->reorder_period > 0 && (bounce % r->reorder_period) == 0;
        }
        if (!sort) {
            tbegin(ST_TRACE, kind);
            WF_TRY(wf_trace(r->scene, rb.ray_buf(), nullptr, rb.hit_buf(), n, rb.closest, ws, nullptr));
            tend();
            return WF_OK;
        }
        const int kb = wf_key_bytes(r->key_mode), bits = wf_key_bits(r->key_mode);
        WF_TRY(rb.prepare_sort(reorder, kb, bits, r->digit_bits));
        if (r->print_sort_logs) {
            // Renderer.cpp:280-283: an unsorted trace first, for comparison.
            tbegin(ST_TRACE_UNSORTED, kind);
            WF_TRY(wf_trace(r->scene, rb.ray_buf(), nullptr, rb.hit_buf(), n, rb.closest, ws, nullptr));
            tend();
        }
        cur_sorted = true;
        tbegin(ST_KEYS, kind);
        WF_TRY(wf_morton_keys(rb.ray_buf(), rb.keys.p, rb.indices.as<uint32_t>(), n, r->key_mode,
                              r->smin, r->smax, ws, nullptr));
        tend();
        tbegin(ST_SORT, kind);
        WF_TRY(wf_sort_pairs(rb.keys.p, rb.indices.as<uint32_t>(), n, kb, bits, r->digit_bits,
                             rb.scratch.p, (int64_t)rb.scratch.bytes, ws, nullptr));
        tend();
        int fuse = 1;
        (void)wf_get_tuning("trace.fuse_reorder", &fuse);
        if (reorder && !fuse) {
            // Tracer::reorderRays then Tracer::trace, timed as two stages
            int o = 1 - rb.swap;
            tbegin(ST_REORDER, kind);
            WF_TRY(wf_reorder(rb.indices.as<uint32_t>(), rb.ray_buf(), rb.rays[o].as<wf_ray>(),
                              rb.s2i_buf(), rb.s2i[o].as<int32_t>(), rb.i2s[o].as<int32_t>(), n,
                              ws, nullptr));
            tend();
            rb.swap = o;
            tbegin(ST_TRACE, kind);
            WF_TRY(wf_trace(r->scene, rb.ray_buf(), nullptr, rb.hit_buf(), n, rb.closest, ws, nullptr));
            tend();
        } else if (reorder) {
            // reorder fused into the trace (its time counts as trace)
            int o = 1 - rb.swap;
            tbegin(ST_TRACE, kind);
            WF_TRY(wf_reorder_trace(r->scene, rb.indices.as<uint32_t>(), rb.ray_buf(),
                                    rb.rays[o].as<wf_ray>(), rb.s2i_buf(), rb.s2i[o].as<int32_t>(),
                                    rb.i2s[o].as<int32_t>(), rb.hit_buf(), n, rb.closest, ws,
                                    nullptr));
            tend();
            rb.swap = o;
        } else {
            tbegin(ST_TRACE, kind);
            WF_TRY(wf_trace(r->scene, rb.ray_buf(), rb.indices.as<uint32_t>(), rb.hit_buf(), n,
                            rb.closest, ws, nullptr));
            tend();
        }
        if (kind == RK_PATH) r->sorted_rays += (uint64_t)n;
        return WF_OK;
    }

    int reconstruct_smooth(RayBuffer& rb, float* px) {
        tbegin(ST_RECONSTRUCT, -1);
        WF_TRY(wf_reconstruct_smooth(r->scene, rb.ray_buf(), rb.hit_buf(), rb.s2i_buf(), px,
                                     r->decreases.as<float>(), rb.size, r->light, r->background,
                                     r->spp, ws, nullptr));
        tend();
        return WF_OK;
    }

    // Renderer::primaryPass (Renderer.cpp:124-131).
    int primary_pass(float* px, int pass) {
        const int64_t np = (int64_t)r->width * r->height;
        RayBuffer& rb = r->primary;
        WF_TRY(rb.resize(np, np));
        rb.closest = true;
        // RayGen::primary: slotToIndex = indexToPixel, indexToSlot = pixelToIndex
        WF_HIP(hipMemcpyAsync(rb.s2i_buf(), r->i2p.p, sizeof(int32_t) * np, hipMemcpyDeviceToDevice, s));
        CameraMatrices cm;
        camera_matrices(r->cam_pos, r->cam_rot, r->fov, (float)r->width / (float)r->height,
                        r->znear, r->zfar, cm);
        tbegin(ST_RAYGEN, RK_PRIMARY);
        WF_TRY(wf_raygen_primary(cm.screen_to_world, r->cam_pos, r->i2p.as<int32_t>(), rb.ray_buf(),
                                 r->width, r->height, pass + r->spp * (r->frame_index - 1), r->zfar,
                                 ws, nullptr));
        tend();
        WF_TRY(trace_rays(rb, RK_PRIMARY, false, false));
        if (r->decision_log)
            WF_TRY(log_hits(rb.hit_buf(), rb.s2i_buf(), np, r->scene->geom_first, dlog_tri(0), s));
        tbegin(ST_RECONSTRUCT, -1);
        WF_TRY(wf_fill_f32(r->decreases.as<float>(), 4 * np, 1.0f, ws, nullptr));
        tend();
        WF_TRY(reconstruct_smooth(rb, px));
        r->n_rays[RK_PRIMARY] += (uint64_t)np;
        return WF_OK;
    }

    // Renderer::shadowPass (Renderer.cpp:133-154).
    int shadow_pass(RayBuffer& in, float* in_px, float* out_px) {
        const int S = r->shadow_samples;
        const int64_t batch = (8ll * 1024 * 1024) / S;
        int64_t begin = 0;
        do {
            int64_t end = std::min(begin + batch, in.size);
            RayBuffer& sh = r->shadow;
            WF_TRY(sh.resize((end - begin) * S, (end - begin) * S));
            sh.closest = false;
            tbegin(ST_RAYGEN, RK_SHADOW);
            WF_TRY(wf_raygen_shadow(r->seeds.as<uint32_t>(), in.ray_buf(), in.hit_buf(), in.s2i_buf(),
                                    sh.ray_buf(), sh.s2i_buf(), sh.i2s_buf(), begin, end, S, r->light,
                                    r->shadow_radius, ws, nullptr));
            tend();
            WF_TRY(trace_rays(sh, RK_SHADOW, r->sort_shadow, r->reorder_shadow));
            if (r->decision_log)
                WF_TRY(log_shadow(sh.hit_buf(), sh.i2s_buf(), in.s2i_buf(), begin, end, S,
                                  dlog_vis(bounce < 0 ? 0 : bounce + 1), s));
            tbegin(ST_RECONSTRUCT, -1);
            WF_TRY(wf_reconstruct_shadow(sh.hit_buf(), in.s2i_buf(), sh.i2s_buf(), in_px, out_px,
                                         begin, end, S, 0, ws, nullptr));
            tend();
            begin = end;
        } while (begin != in.size);
        // Renderer::computeRayHits: shadow rays counted as S x hits.
        WF_TRY(wf_count_hits(in.hit_buf(), in.size, r->counters.as<uint32_t>(), ws, nullptr));
        return WF_OK;
    }

    // Renderer::renderPath (Renderer.cpp:175-199).
    int render_path(int pass_index) {
        const int64_t np = (int64_t)r->width * r->height;
        pass = pass_index;
        bounce = -1;
        float* aux = r->aux_pixels.as<float>();
        float* fp = r->frame_pixels.as<float>();
        if (r->ray_type == "primary") return primary_pass(fp, pass);  // renderPrimary
        WF_TRY(fill_u32(r->aux_pixels.as<uint32_t>(), 4 * np, 0u, s));
        WF_TRY(primary_pass(aux, pass));
        WF_TRY(shadow_pass(r->primary, aux, fp));
        if (r->ray_type == "shadow") return WF_OK;
        int64_t cap = r->primary.size;
        for (int q = 0; q < 2; ++q) WF_TRY(r->queue[q].resize(cap, np));
        WF_TRY(r->compact.ensure((size_t)wf_compact_scratch_bytes(cap)));
        for (bounce = 0; bounce < r->depth; ++bounce) {
            WF_TRY(fill_u32(r->aux_pixels.as<uint32_t>(), 4 * np, 0u, s));
            RayBuffer& in = bounce > 0 ? r->queue[r->queue_swap] : r->primary;
            RayBuffer& out = r->queue[1 - r->queue_swap];
            // RayGen::path (RayGen.cpp:135-186)
            WF_TRY(out.resize(in.size, np));
            out.closest = true;
            uint32_t count = 0;
            tbegin(ST_SHADE, RK_PATH);
            WF_TRY(shade_extend_logged(r->scene, r->seeds.as<uint32_t>(), in.ray_buf(), in.hit_buf(),
                                       in.s2i_buf(), in.size, out.ray_buf(), out.s2i_buf(),
                                       r->decreases.as<float>(), r->counters.as<uint32_t>() + 1,
                                       r->russian_roulette ? 1 : 0, bounce, r->rr_start,
                                       r->compact.p, (int64_t)r->compact.bytes, nullptr, ws, nullptr,
                                       r->decision_log ? dlog_sign(bounce) : nullptr));
            tend();
            WF_HIP(hipMemcpyAsync(&count, r->counters.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, s));
            WF_HIP(hipStreamSynchronize(s));
            out.size = count;
            if (r->capture && count > 0) {
                // slot = pass * recursionDepth + bounce
                const size_t slot = (size_t)pass * (size_t)r->depth + (size_t)bounce;
                while (r->cap.size() <= slot) r->cap.emplace_back(new DeviceBuffer());
                if (r->cap_n.size() <= slot) r->cap_n.resize(slot + 1, 0);
                // rays followed by the slot-to-pixel LUT
                WF_TRY(r->cap[slot]->ensure((sizeof(wf_ray) + sizeof(int32_t)) * count));
                WF_HIP(hipMemcpyAsync(r->cap[slot]->p, out.ray_buf(), sizeof(wf_ray) * count,
                                      hipMemcpyDeviceToDevice, s));
                WF_HIP(hipMemcpyAsync(r->cap[slot]->as<char>() + sizeof(wf_ray) * count,
                                      out.s2i_buf(), sizeof(int32_t) * count,
                                      hipMemcpyDeviceToDevice, s));
                r->cap_n[slot] = count;
            }
            WF_TRY(trace_rays(out, RK_PATH, r->sort_path, r->reorder_path));
            if (r->decision_log)
                WF_TRY(log_hits(out.hit_buf(), out.s2i_buf(), count, r->scene->geom_first,
                                dlog_tri(bounce + 1), s));
            WF_TRY(reconstruct_smooth(out, aux));
            r->n_rays[RK_PATH] += count;
            if (count == 0) break;
            WF_TRY(shadow_pass(out, aux, fp));
            r->queue_swap = 1 - r->queue_swap;
        }
        return WF_OK;
    }
};

// Renderer::render (Renderer.cpp:552-654).
static int render_frame(wf_renderer_s* r, float* ms_out) {
    WF_CHECK_ARG(r->scene, "wf_renderer_render: no scene (call wf_renderer_build_scene)");
    WF_TRY(ensure_stream(r));
    hipStream_t s = r->stream;
    const int64_t np = (int64_t)r->width * r->height;
    WF_TRY(r->pixels.ensure(16 * np));
    WF_TRY(r->frame_pixels.ensure(16 * np));
    WF_TRY(r->aux_pixels.ensure(16 * np));
    WF_TRY(r->decreases.ensure(16 * np));
    WF_TRY(r->seeds.ensure(4 * np));
    WF_TRY(r->counters.ensure(64));
    if (r->lut_w != r->width || r->lut_h != r->height) {
        std::vector<int32_t> i2p, p2i;
        pixel_table(r->width, r->height, i2p, p2i);
        WF_TRY(r->i2p.ensure(4 * np));
        WF_HIP(hipMemcpy(r->i2p.p, i2p.data(), 4 * np, hipMemcpyHostToDevice));
        r->lut_w = r->width;
        r->lut_h = r->height;
    }
    WF_TRY(fill_u32(r->frame_pixels.as<uint32_t>(), 4 * np, 0u, s));
    if (r->frame_index == 1) WF_TRY(fill_u32(r->pixels.as<uint32_t>(), 4 * np, 0u, s));
    WF_TRY(fill_u32(r->counters.as<uint32_t>(), 16, 0u, s));
    for (int k = 0; k < 3; ++k) {
        r->n_rays[k] = 0;
        r->trace_ms[k] = 0.f;
    }
    r->sorted_rays = 0;
    r->log.reset();
    r->cap_n.assign(r->cap_n.size(), 0);
    if (r->decision_log) {
        const int64_t slots = (int64_t)r->spp * (r->depth + 1) * np;
        const int64_t vis_words = (slots * r->shadow_samples + 3) / 4;
        WF_TRY(r->dlog_tri.ensure(4 * slots));
        WF_TRY(r->dlog_vis.ensure(4 * vis_words));
        WF_TRY(fill_u32(r->dlog_tri.as<uint32_t>(), slots, 0xFFFFFFFEu, s));   // -2: no ray
        WF_TRY(fill_u32(r->dlog_vis.as<uint32_t>(), vis_words, 0xFFFFFFFFu, s));
        WF_TRY(r->dlog_sign.ensure(4 * ((slots + 3) / 4)));
        WF_TRY(fill_u32(r->dlog_sign.as<uint32_t>(), (slots + 3) / 4, 0xFFFFFFFFu, s));
    }
    Frame f{r, s, reinterpret_cast<wf_stream_t>(s)};
    if (r->ray_type != "primary") {
        f.tbegin(ST_RAYGEN, -1);
        WF_TRY(wf_init_seeds(r->seeds.as<uint32_t>(), np, r->frame_index, f.ws, nullptr));
        f.tend();
    }
    for (int pass = 0; pass < r->spp; ++pass) WF_TRY(f.render_path(pass));
    f.tbegin(ST_RECONSTRUCT, -1);
    WF_TRY(wf_interpolate_colors(r->frame_pixels.as<float>(), r->pixels.as<float>(), np,
                                 r->frame_index, r->key_value, r->white_point, f.ws, nullptr));
    f.tend();
    ++r->frame_index;
    uint32_t hits = 0;
    WF_HIP(hipMemcpyAsync(&hits, r->counters.p, 4, hipMemcpyDeviceToHost, s));
    WF_HIP(hipStreamSynchronize(s));
    r->n_rays[RK_SHADOW] = (uint64_t)r->shadow_samples * hits;
    for (int k = 0; k < 8; ++k) r->stage_ms[k] = 0.f;
    r->path_logs.assign((size_t)r->depth, wf_sort_log{});
    r->shadow_logs.assign((size_t)r->depth + 1, wf_sort_log{});
    float total = 0.f;
    for (auto& rec : r->log.recs) {
        float t = 0.f;
        WF_HIP(hipEventElapsedTime(&t, rec.a, rec.b));
        total += t;
        // SortLog (Renderer.cpp:280-289): path logs by bounce, shadow logs by
        // the vertex the shadow rays leave from (0 = primary hit)
        wf_sort_log* sl = nullptr;
        if (rec.kind == RK_PATH && rec.bounce >= 0 && rec.bounce < r->depth)
            sl = &r->path_logs[(size_t)rec.bounce];
        if (rec.kind == RK_SHADOW && rec.bounce + 1 >= 0 && rec.bounce + 1 <= r->depth)
            sl = &r->shadow_logs[(size_t)(rec.bounce + 1)];
        if (sl) {
            switch (rec.stage) {
                case ST_KEYS: sl->morton_ms += t; sl->sorted_rays += rec.n; break;
                case ST_SORT: sl->sort_ms += t; break;
                case ST_REORDER: sl->reorder_ms += t; break;
                case ST_TRACE_UNSORTED: sl->trace_ms += t; break;
                case ST_TRACE:
                    sl->ray_count += rec.n;
                    (rec.sorted ? sl->trace_sort_ms : sl->trace_ms) += t;
                    break;
                default: break;
            }
        }
        if (rec.kind == RK_PATH || rec.stage == ST_TRACE_UNSORTED) {
            // stage_ms: {keys, sort, reorder, trace_sorted, trace_unsorted, shade, raygen, reconstruct}
            if (rec.kind == RK_PATH) {
                switch (rec.stage) {
                    case ST_KEYS: r->stage_ms[0] += t; break;
                    case ST_SORT: r->stage_ms[1] += t; break;
                    case ST_REORDER: r->stage_ms[2] += t; break;
                    case ST_TRACE: r->stage_ms[3] += t; break;
                    case ST_TRACE_UNSORTED: r->stage_ms[4] += t; break;
                    case ST_SHADE: r->stage_ms[5] += t; break;
                    default: break;
                }
            }
        }
        if (rec.stage == ST_RAYGEN) r->stage_ms[6] += t;
        if (rec.stage == ST_RECONSTRUCT) r->stage_ms[7] += t;
        // Renderer trace times: raygen (shade for paths) + keys + sort + reorder + trace
        if (rec.kind >= 0 && rec.stage != ST_TRACE_UNSORTED) r->trace_ms[rec.kind] += t;
    }
    if (ms_out) *ms_out = total;
    return WF_OK;
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_renderer_create(const char* env_json_path, wf_renderer* out) {
    WF_CHECK_ARG(env_json_path, "wf_renderer_create: null path");
    return create(env_json_path, true, out);
}
extern "C" int wf_renderer_create_from_string(const char* env_json_text, wf_renderer* out) {
    WF_CHECK_ARG(env_json_text, "wf_renderer_create_from_string: null text");
    return create(env_json_text, false, out);
}
extern "C" int wf_renderer_destroy(wf_renderer r) {
    delete r;
    return WF_OK;
}

// The scene of Scene.procedural (this library's generators) or of
// Scene.filename (a glTF file, VulkanExample::prepare,
// WaveFrontPathTracer.cpp:505-535), uploaded and bound with its bounds.
extern "C" int wf_renderer_build_scene(wf_renderer r) {
    WF_CHECK_ARG(r, "wf_renderer_build_scene: null renderer");
    std::string kind, file;
    HostScene hs;
    if (r->env.get_string("Scene.procedural", kind) && !kind.empty()) {
        int k = kind == "boxes" ? WF_SCENE_BOXES
                : kind == "interior" ? WF_SCENE_INTERIOR
                : kind == "city" ? WF_SCENE_CITY : -1;
        WF_CHECK_ARG(k >= 0, "Scene.procedural must be boxes, interior or city");
        int seed = 1000;
        r->env.get_int("Scene.seed", seed);
        WF_TRY(procedural_scene(k, (uint64_t)seed, hs));
    } else if (r->env.get_string("Scene.filename", file) && !file.empty()) {
        GltfSizes sz;
        WF_TRY(load_gltf(file[0] == '/' ? file : r->asset_path + file, &hs, sz));
    } else {
        return fail(WF_ERR_INVALID_ARG, "neither Scene.procedural nor Scene.filename is set");
    }
    WF_TRY(ensure_stream(r));
    int groups = 0;
    r->env.get_int("Scene.materialGroups", groups);
    if (groups > 0) WF_TRY(regroup_materials(hs, groups));
    const int64_t nv = (int64_t)hs.vertices.size() / 6, nt = (int64_t)hs.indices.size() / 3;
    DeviceBuffer v, idx, mat;
    WF_TRY(v.ensure(sizeof(float) * hs.vertices.size()));
    WF_TRY(idx.ensure(sizeof(uint32_t) * hs.indices.size()));
    WF_TRY(mat.ensure(sizeof(int32_t) * hs.tri_material.size()));
    WF_HIP(hipMemcpy(v.p, hs.vertices.data(), sizeof(float) * hs.vertices.size(), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(idx.p, hs.indices.data(), sizeof(uint32_t) * hs.indices.size(), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(mat.p, hs.tri_material.data(), sizeof(int32_t) * hs.tri_material.size(), hipMemcpyHostToDevice));
    wf_scene sc = nullptr;
    wf_stream_t ws = reinterpret_cast<wf_stream_t>(r->stream);
    // Scene.splitDepth: the presplit depth of this scene's hierarchy (the
    // process-wide bvh.split_depth is restored after the build)
    int split = -1, split_saved = 0;
    r->env.get_int("Scene.splitDepth", split);
    if (split >= 0) {
        WF_TRY(wf_get_tuning("bvh.split_depth", &split_saved));
        WF_TRY(wf_set_tuning("bvh.split_depth", split));
    }
    const int built = wf_scene_create(v.p, nv, 24, idx.as<uint32_t>(), nt, hs.geom_first.data(),
                                      (int)hs.geom_first.size() - 1, mat.as<int32_t>(), ws, &sc);
    if (split >= 0) (void)wf_set_tuning("bvh.split_depth", split_saved);  // a value read back
    WF_TRY(built);
    int st = wf_scene_set_materials(sc, hs.albedo.data(), hs.emission.data(),
                                    (int)hs.albedo.size() / 3, ws);
    if (st != WF_OK) {
        wf_scene_destroy(sc);
        return st;
    }
    if (r->own_scene && r->scene) wf_scene_destroy(r->scene);
    r->scene = sc;
    r->own_scene = true;
    // Renderer::setSceneBounds with the model's dimensions
    for (int a = 0; a < 3; ++a) {
        r->smin[a] = hs.bmin[a];
        r->smax[a] = hs.bmax[a];
    }
    r->frame_index = 1;
    return WF_OK;
}

extern "C" int wf_renderer_set_asset_path(wf_renderer r, const char* dir) {
    WF_CHECK_ARG(r && dir, "wf_renderer_set_asset_path: null argument");
    r->asset_path = dir;
    if (!r->asset_path.empty() && r->asset_path.back() != '/') r->asset_path += '/';
    return WF_OK;
}

extern "C" int wf_renderer_set_scene(wf_renderer r, wf_scene scene, const float scene_min[3],
                                     const float scene_max[3]) {
    WF_CHECK_ARG(r && scene, "wf_renderer_set_scene: null argument");
    if (r->own_scene && r->scene && r->scene != scene) wf_scene_destroy(r->scene);
    r->scene = scene;
    r->own_scene = false;
    if (scene_min && scene_max) {
        for (int a = 0; a < 3; ++a) {
            r->smin[a] = scene_min[a];
            r->smax[a] = scene_max[a];
        }
    } else {
        wf_scene_bounds(scene, r->smin, r->smax);
    }
    r->frame_index = 1;
    return WF_OK;
}

extern "C" int wf_renderer_get_scene(wf_renderer r, wf_scene* out) {
    WF_CHECK_ARG(r && out, "wf_renderer_get_scene: null argument");
    *out = r->scene;
    return WF_OK;
}

extern "C" int wf_renderer_render(wf_renderer r, float* ms_out) {
    WF_CHECK_ARG(r, "wf_renderer_render: null renderer");
    return render_frame(r, ms_out);
}

extern "C" int wf_renderer_reset_frame_index(wf_renderer r) {
    WF_CHECK_ARG(r, "wf_renderer_reset_frame_index: null renderer");
    r->frame_index = 1;
    return WF_OK;
}

extern "C" int wf_renderer_pixels_device(wf_renderer r, int tone_mapped, const float** out) {
    WF_CHECK_ARG(r && out, "wf_renderer_pixels_device: null argument");
    const DeviceBuffer& b = tone_mapped ? r->frame_pixels : r->pixels;
    WF_CHECK_ARG(b.p && r->stream, "wf_renderer_pixels_device: nothing rendered yet");
    WF_HIP(hipStreamSynchronize(r->stream));
    *out = b.as<float>();
    return WF_OK;
}

// Path-replay adjoint of the last rendered frame (csrc/adjoint.hip).
extern "C" int wf_renderer_adjoint(wf_renderer r, int frame, const float* d_pixels,
                                   float* grad_albedo, float* grad_emission, float* replay,
                                   float* ms) {
    WF_CHECK_ARG(r && d_pixels && grad_albedo && grad_emission,
                 "wf_renderer_adjoint: null argument");
    WF_CHECK_ARG(r->scene, "wf_renderer_adjoint: no scene");
    WF_CHECK_ARG(r->frame_index > 1 && r->lut_w == r->width && r->lut_h == r->height,
                 "wf_renderer_adjoint: render a frame first");
    WF_CHECK_ARG(frame >= 0, "wf_renderer_adjoint: frame < 0");
    if (frame == 0) frame = r->frame_index - 1;
    WF_CHECK_ARG(r->ray_type == "path", "wf_renderer_adjoint: needs Renderer.rayType = path");
    WF_CHECK_ARG(!r->russian_roulette, "wf_renderer_adjoint: Russian roulette is not differentiable here");
    WF_CHECK_ARG(r->depth >= 0 && r->depth <= ADJ_MAX_VERTS,
                 "wf_renderer_adjoint: recursionDepth must be <= 8");
    WF_TRY(ensure_stream(r));
    hipStream_t s = r->stream;
    wf_scene sc = r->scene;
    const int nmat = sc->num_materials;
    WF_TRY(fill_u32(reinterpret_cast<uint32_t*>(grad_albedo), 3 * (int64_t)nmat, 0u, s));
    WF_TRY(fill_u32(reinterpret_cast<uint32_t*>(grad_emission), 3 * (int64_t)nmat, 0u, s));
    AdjointParams p{};
    p.nodes = sc->nodes;
    p.nodes4 = sc->nodes4;
    p.tris = sc->tris;
    p.shade = sc->shade;
    p.albedo = sc->albedo;
    p.emission = sc->emission;
    p.nmat = nmat;
    CameraMatrices cm;
    camera_matrices(r->cam_pos, r->cam_rot, r->fov, (float)r->width / (float)r->height, r->znear,
                    r->zfar, cm);
    for (int k = 0; k < 16; ++k) p.s2w.m[k] = cm.screen_to_world[k];
    for (int a = 0; a < 3; ++a) {
        p.origin[a] = r->cam_pos[a];
        p.light[a] = r->light[a];
        p.background[a] = r->background[a];
    }
    p.index_to_pixel = r->i2p.as<int32_t>();
    p.width = r->width;
    p.height = r->height;
    p.spp = r->spp;
    p.depth = r->depth;
    p.frame = frame;
    p.S = r->shadow_samples;
    p.radius = r->shadow_radius;
    p.max_dist = r->zfar;
    p.d_pixels = reinterpret_cast<const float4*>(d_pixels);
    p.grad_albedo = grad_albedo;
    p.grad_emission = grad_emission;
    p.replay = reinterpret_cast<float4*>(replay);
    StageTimer t(s, ms);
    WF_TRY(adjoint_launch(p, s));
    WF_TRY(t.finish());
    WF_HIP(hipStreamSynchronize(s));
    return WF_OK;
}

extern "C" int wf_renderer_save_image(wf_renderer r, const char* path, int linear) {
    WF_CHECK_ARG(r && path, "wf_renderer_save_image: null argument");
    std::vector<float> px(4 * (size_t)r->width * r->height);
    WF_TRY(wf_renderer_read_pixels(r, linear ? 0 : 1, px.data()));
    return linear ? wf_write_pfm(path, px.data(), r->width, r->height)
                  : wf_write_png(path, px.data(), r->width, r->height);
}

extern "C" int wf_renderer_read_pixels(wf_renderer r, int tone_mapped, float* host_rgba) {
    WF_CHECK_ARG(r && host_rgba, "wf_renderer_read_pixels: null argument");
    const DeviceBuffer& b = tone_mapped ? r->frame_pixels : r->pixels;
    WF_CHECK_ARG(b.p && r->stream, "wf_renderer_read_pixels: nothing rendered yet");
    WF_HIP(hipStreamSynchronize(r->stream));
    WF_HIP(hipMemcpy(host_rgba, b.p, 16 * (size_t)r->width * r->height, hipMemcpyDeviceToHost));
    return WF_OK;
}

extern "C" int wf_renderer_size(wf_renderer r, int* width, int* height) {
    WF_CHECK_ARG(r, "wf_renderer_size: null renderer");
    if (width) *width = r->width;
    if (height) *height = r->height;
    return WF_OK;
}

extern "C" int wf_renderer_stats(wf_renderer r, uint64_t counts[3], float times[3],
                                 float stage_ms[8], uint64_t* sorted_rays) {
    WF_CHECK_ARG(r, "wf_renderer_stats: null renderer");
    for (int k = 0; k < 3; ++k) {
        if (counts) counts[k] = r->n_rays[k];
        if (times) times[k] = r->trace_ms[k];
    }
    if (stage_ms)
        for (int k = 0; k < 8; ++k) stage_ms[k] = r->stage_ms[k];
    if (sorted_rays) *sorted_rays = r->sorted_rays;
    return WF_OK;
}

extern "C" int wf_renderer_sort_logs(wf_renderer r, int kind, wf_sort_log* out, int max_entries,
                                     int* num_entries) {
    WF_CHECK_ARG(r, "wf_renderer_sort_logs: null renderer");
    WF_CHECK_ARG(kind == 0 || kind == 1, "wf_renderer_sort_logs: kind is 0 (path) or 1 (shadow)");
    const std::vector<wf_sort_log>& v = kind == 0 ? r->path_logs : r->shadow_logs;
    if (num_entries) *num_entries = (int)v.size();
    if (out)
        for (int k = 0; k < std::min(max_entries, (int)v.size()); ++k) out[k] = v[(size_t)k];
    return WF_OK;
}

extern "C" int wf_renderer_set_decision_log(wf_renderer r, int enable) {
    WF_CHECK_ARG(r, "wf_renderer_set_decision_log: null renderer");
    r->decision_log = enable != 0;
    return WF_OK;
}

extern "C" int wf_renderer_decisions(wf_renderer r, int32_t* hit_tri, uint8_t* shadow_visible,
                                     uint8_t* bounce_signs) {
    WF_CHECK_ARG(r, "wf_renderer_decisions: null renderer");
    if (!r->decision_log || !r->dlog_tri.p)
        return fail(WF_ERR_STATE, "wf_renderer_decisions: render a frame with the decision log on first");
    const int64_t slots = (int64_t)r->spp * (r->depth + 1) * r->width * r->height;
    WF_HIP(hipStreamSynchronize(r->stream));
    if (hit_tri) WF_HIP(hipMemcpy(hit_tri, r->dlog_tri.p, 4 * slots, hipMemcpyDeviceToHost));
    if (shadow_visible)
        WF_HIP(hipMemcpy(shadow_visible, r->dlog_vis.p, (size_t)slots * r->shadow_samples,
                         hipMemcpyDeviceToHost));
    if (bounce_signs) WF_HIP(hipMemcpy(bounce_signs, r->dlog_sign.p, (size_t)slots, hipMemcpyDeviceToHost));
    return WF_OK;
}

extern "C" int wf_env_get_int(wf_renderer r, const char* name, int* value) {
    WF_CHECK_ARG(r && name && value, "wf_env_get_int: null argument");
    return r->env.get_int(name, *value) ? WF_OK : fail(WF_ERR_INVALID_ARG, std::string("no value for ") + name);
}
extern "C" int wf_env_get_float(wf_renderer r, const char* name, float* value) {
    WF_CHECK_ARG(r && name && value, "wf_env_get_float: null argument");
    return r->env.get_float(name, *value) ? WF_OK : fail(WF_ERR_INVALID_ARG, std::string("no value for ") + name);
}
extern "C" int wf_env_get_string(wf_renderer r, const char* name, char* buf, int buf_len) {
    WF_CHECK_ARG(r && name && buf && buf_len > 0, "wf_env_get_string: null argument");
    std::string v;
    if (!r->env.get_string(name, v)) return fail(WF_ERR_INVALID_ARG, std::string("no value for ") + name);
    std::strncpy(buf, v.c_str(), (size_t)buf_len - 1);
    buf[buf_len - 1] = 0;
    return WF_OK;
}
extern "C" int wf_env_get_vector(wf_renderer r, const char* name, float value[3]) {
    WF_CHECK_ARG(r && name && value, "wf_env_get_vector: null argument");
    return r->env.get_vector(name, value) ? WF_OK : fail(WF_ERR_INVALID_ARG, std::string("no value for ") + name);
}

extern "C" int wf_renderer_set_capture(wf_renderer r, int enable) {
    WF_CHECK_ARG(r, "wf_renderer_set_capture: null renderer");
    r->capture = enable != 0;
    return WF_OK;
}

extern "C" int wf_renderer_captured(wf_renderer r, int bounce, const wf_ray** rays,
                                    const int32_t** slot_to_index, int64_t* n) {
    WF_CHECK_ARG(r && rays && n, "wf_renderer_captured: null argument");
    if (bounce < 0 || bounce >= (int)r->cap_n.size() || r->cap_n[bounce] == 0) {
        *rays = nullptr;
        if (slot_to_index) *slot_to_index = nullptr;
        *n = 0;
        return WF_OK;
    }
    *rays = r->cap[bounce]->as<wf_ray>();
    if (slot_to_index)
        *slot_to_index = reinterpret_cast<const int32_t*>(r->cap[bounce]->as<char>() +
                                                          sizeof(wf_ray) * r->cap_n[bounce]);
    *n = r->cap_n[bounce];
    return WF_OK;
}

extern "C" int wf_trace_sort(wf_scene scene, wf_ray_buffer* rb, int key_mode, int digit_bits,
                             const float scene_min[3], const float scene_max[3], int reorder,
                             wf_stream_t stream, float times[4]) {
    WF_CHECK_ARG(scene && rb && scene_min && scene_max, "wf_trace_sort: null argument");
    const int kb = wf_key_bytes(key_mode), bits = wf_key_bits(key_mode);
    WF_CHECK_ARG(kb > 0 && key_mode != WF_KEY_SCENE6_ORIGIN30_OCT28, "wf_trace_sort: bad key mode");
    const int64_t n = rb->size;
    WF_CHECK_ARG(n >= 0 && n <= rb->capacity, "wf_trace_sort: size outside capacity");
    float t[4] = {0, 0, 0, 0};
    float* tp[4] = {times ? &t[0] : nullptr, times ? &t[1] : nullptr, times ? &t[2] : nullptr,
                    times ? &t[3] : nullptr};
    const int cur = rb->swap ? 1 : 0, o = 1 - cur;
    WF_CHECK_ARG(rb->rays[cur] && rb->results && rb->keys && rb->indices && rb->sort_scratch,
                 "wf_trace_sort: missing buffer");
    WF_TRY(wf_morton_keys(rb->rays[cur], rb->keys, rb->indices, n, key_mode, scene_min, scene_max,
                          stream, tp[0]));
    WF_TRY(wf_sort_pairs(rb->keys, rb->indices, n, kb, bits, digit_bits, rb->sort_scratch,
                         rb->sort_scratch_bytes, stream, tp[1]));
    if (reorder) {
        WF_CHECK_ARG(rb->rays[o] && rb->slot_to_index[cur] && rb->slot_to_index[o] && rb->index_to_slot[o],
                     "wf_trace_sort: reorder buffers missing");
        // reorder and trace fused: the time is reported as trace, reorder 0
        WF_TRY(wf_reorder_trace(scene, rb->indices, rb->rays[cur], rb->rays[o], rb->slot_to_index[cur],
                                rb->slot_to_index[o], rb->index_to_slot[o], rb->results, n,
                                rb->closest_hit, stream, tp[3]));
        rb->swap = o;
    } else {
        WF_TRY(wf_trace(scene, rb->rays[cur], rb->indices, rb->results, n, rb->closest_hit, stream,
                        tp[3]));
    }
    if (times)
        for (int k = 0; k < 4; ++k) times[k] = t[k];
    return WF_OK;
}
