/*
 * wfpt.h - C ABI of the MI355X wavefront path tracer.
 *
 * Every entry point is stream-ordered on the hipStream_t passed in (NULL = the
 * default stream) and returns an int status (WF_OK = 0, negative = error; see
 * wf_status_string).  Buffers are device pointers owned by the caller unless a
 * function says otherwise.  When a `float* ms` out-parameter is non-NULL the
 * call brackets its launches with hipEvents, waits for them and writes the GPU
 * time in milliseconds, which is what the reference's ComputePass::launchTimed
 * returns (projects/WaveFrontPathTracer/Compute/ComputePass.cpp:85-110); with
 * ms == NULL the call only enqueues work.
 *
 * Record layouts are byte-compatible with the reference's
 * Ray / RayResult (projects/WaveFrontPathTracer/Ray/Ray.h:32-44).
 */
#ifndef WFPT_H
#define WFPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* wf_stream_t; /* == hipStream_t */

/* ---------------------------------------------------------------- records */

/* Ray.h:32-37 : {origin.xyz, tmin, direction.xyz, tmax}, 32 bytes. */
typedef struct wf_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} wf_ray;

/* Ray.h:39-44 : {primitiveId, geometryId, t, packUnorm2x16(bary)}, 16 bytes.
 * A miss is {-1, -1, -1.0f, 0} (shaders/glsl/WaveFrontPathTracer/trace.comp:78-81). */
typedef struct wf_hit {
    int32_t prim;
    int32_t geom;
    float t;
    uint32_t bary;
} wf_hit;

/* --------------------------------------------------------------- statuses */
enum {
    WF_OK = 0,
    WF_ERR_INVALID_ARG = -1,  /* null pointer, bad size, unknown mode      */
    WF_ERR_HIP = -2,          /* a HIP runtime call failed                  */
    WF_ERR_OUT_OF_MEMORY = -3,
    WF_ERR_SCRATCH_TOO_SMALL = -4,
    WF_ERR_TOO_LARGE = -5,    /* element count beyond the kernels' limits   */
    WF_ERR_PARSE = -6,        /* malformed env.json                         */
    WF_ERR_IO = -7,           /* file cannot be opened                      */
    WF_ERR_STATE = -8         /* object used before it was set up           */
};
const char* wf_status_string(int status);
/* Last error message of the calling thread (empty string if none). */
const char* wf_last_error(void);
/* Library version string. */
const char* wf_version(void);

/* Largest element count accepted by the sort / compaction kernels (2^30-1):
 * look-back words carry a 30-bit count. */
#define WF_MAX_ELEMENTS ((int64_t)((1u << 30) - 1u))

/* ------------------------------------------------------------ Morton keys */
/* Replaces Tracer::computeMortonCodes (Tracer/Tracer.cpp:22-37) and
 * computeMortonCodesAila32.comp:24-96.
 *   WF_KEY_AILA32          : the reference's 32-bit key (7 origin bits/axis
 *                            interleaved with 4 direction bits/axis), u32.
 *   WF_KEY_ORIGIN30        : 30-bit origin Morton code (10 bits/axis), u32.
 *   WF_KEY_ORIGIN30_OCT30  : 60-bit key = origin30 << 30 | 2-D Morton code
 *                            of the 15+15-bit octahedral direction, u64.
 *   WF_KEY_SCENE6_ORIGIN30_OCT28 : batched key = scene_id << 58 |
 *                            origin30 << 28 | 14+14-bit octahedral, u64.
 *   WF_KEY_ORIGIN_DIR_MORTON60 : the 30 origin and 30 direction bits of
 *                            ORIGIN30_OCT30 as one 5-D Morton code: from the
 *                            top, 10 levels of (u, v, x, y, z) bits (direction
 *                            bits 14..5 with origin bits 9..0), then the
 *                            direction's low 5 levels as (u, v) pairs, u64;
 *                            in the spirit of the reference's interleaved
 *                            Aila32 key (origin and direction bits alternate).
 *   WF_KEY_ORIGIN_DIR_MORTON30 : the top 30 bits of ORIGIN_DIR_MORTON60 (6
 *                            levels of (u, v, x, y, z): direction bits 14..9
 *                            with origin bits 9..4), u32: the same coherence
 *                            at half the sort passes.  No counterpart in the
 *                            reference.
 * keys receive u32 (AILA32, ORIGIN30, ORIGIN_DIR_MORTON30) or u64 (others);
 * indices[i] = i. */
enum {
    WF_KEY_AILA32 = 0,
    WF_KEY_ORIGIN30 = 1,
    WF_KEY_ORIGIN30_OCT30 = 2,
    WF_KEY_SCENE6_ORIGIN30_OCT28 = 3,
    WF_KEY_ORIGIN_DIR_MORTON60 = 4,
    WF_KEY_ORIGIN_DIR_MORTON30 = 5
};
int wf_key_bytes(int mode);  /* 4 or 8, or WF_ERR_INVALID_ARG */
int wf_key_bits(int mode);   /* significant bits: 32, 30, 60, 64, 60, 30 */
int wf_morton_keys(const wf_ray* rays, void* keys, uint32_t* indices, int64_t n,
                   int mode, const float scene_min[3], const float scene_max[3],
                   wf_stream_t stream, float* ms);
/* Batched variant: scene_of_ray[i] selects scene_min/max[3*scene] and the
 * 6-bit prefix (scene < 64). */
int wf_morton_keys_batched(const wf_ray* rays, const uint8_t* scene_of_ray,
                           void* keys, uint32_t* indices, int64_t n,
                           const float* scene_min, const float* scene_max,
                           int num_scenes, wf_stream_t stream, float* ms);

/* -------------------------------------------------------------- radix sort */
/* Replaces vrdxCmdSortKeyValue (external/vulkan_radix_sort/vk_radix_sort.h:
 * 2579-2585, gpuSort :2598-2755) as called by Tracer::radixSort
 * (Tracer/Tracer.cpp:39-78): stable ascending LSD sort of keys[n] (u32 when
 * key_bytes == 4, u64 when 8) carrying vals[n] (u32), result in place.
 * Only the low key_bits bits take part; passes = ceil(key_bits/digit_bits),
 * digit_bits in {4, 8, 11}.  scratch must hold wf_sort_scratch_bytes(...). */
int64_t wf_sort_scratch_bytes(int64_t n, int key_bytes, int key_bits, int digit_bits);
int wf_sort_pairs(void* keys, uint32_t* vals, int64_t n, int key_bytes, int key_bits,
                  int digit_bits, void* scratch, int64_t scratch_bytes,
                  wf_stream_t stream, float* ms);
/* As wf_sort_pairs, with the element count read on the device from
 * *d_count (clamped to capacity; scratch sized for capacity), so a captured
 * bounce loop can sort the compacted wavefront without the host readback of
 * the count that RayGen::path makes (RayGen/RayGen.cpp:178-183). */
int wf_sort_pairs_count(void* keys, uint32_t* vals, int64_t capacity,
                        const uint32_t* d_count, int key_bytes, int key_bits,
                        int digit_bits, void* scratch, int64_t scratch_bytes,
                        wf_stream_t stream);
/* No counterpart in the reference: wf_sort_pairs for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The scratch is a 32 MiB device array inside the
 * library, one per GPU, held until the process ends; the call allocates
 * nothing and only enqueues work (no timing).  n is at most
 * wf_sort_noalloc_max_pairs() = 1,048,576 pairs (WF_ERR_TOO_LARGE above).
 * All calls on one GPU share its array: calls that overlap on one GPU
 * (different streams or threads) race and must be serialised by the caller. */
int64_t wf_sort_noalloc_max_pairs(void);
int wf_sort_pairs_noalloc(void* keys, uint32_t* vals, int64_t n, int key_bytes, int key_bits,
                          int digit_bits, wf_stream_t stream);

/* ----------------------------------------------------------------- reorder */
/* Replaces Tracer::reorderRays (Tracer/Tracer.cpp:80-98), reorderRays.comp:
 * out_rays[i] = in_rays[idx[i]]; out_s2i[i] = in_s2i[idx[i]];
 * out_i2s[out_s2i[i]] = i.  out_i2s must hold max(in_s2i)+1 entries. */
int wf_reorder(const uint32_t* idx, const wf_ray* in_rays, wf_ray* out_rays,
               const int32_t* in_s2i, int32_t* out_s2i, int32_t* out_i2s, int64_t n,
               wf_stream_t stream, float* ms);

/* ------------------------------------------------------------------- scene */
/* Replaces the BLAS/TLAS build (WaveFrontPathTracer.cpp:202-350, :355-457)
 * and Tracer::setAccererationStructure (Tracer/Tracer.cpp:166-183) with a
 * hierarchy built on the GPU (binned SAH by default, bvh.builder).  Input
 * (device): vertex records of vertex_stride bytes holding position float3
 * at byte 0 and normal float3 at byte 12 (vkglTF::Vertex,
 * base/VulkanglTFModel.h:221-228 uses stride 96), and a u32 index buffer.
 * geom_first_tri (host, num_geoms+1 entries, ascending, geom_first_tri[0] = 0)
 * splits the triangles into geometries: a hit reports
 * prim = tri - geom_first_tri[geom] and geom, as the reference's
 * gl_PrimitiveID / gl_GeometryIndex (trace.comp:71-72).
 * tri_material (device, optional) gives each triangle a material index
 * (default 0).  The scene keeps its own copies; inputs may be freed after. */
typedef struct wf_scene_s* wf_scene;
int wf_scene_create(const void* vertices, int64_t num_vertices, int vertex_stride,
                    const uint32_t* indices, int64_t num_tris,
                    const int32_t* geom_first_tri, int num_geoms,
                    const int32_t* tri_material, wf_stream_t stream, wf_scene* out);
int wf_scene_destroy(wf_scene scene);
/* Per-material albedo (rgb) and emission (rgb): host or device arrays of
 * 3*num_materials floats, copied (the device copies are reused while the
 * count stays the same). Default one material: albedo 0.8
 * (raygenPath.comp:138), emission 0. */
int wf_scene_set_materials(wf_scene scene, const float* albedo_rgb,
                           const float* emission_rgb, int num_materials,
                           wf_stream_t stream);
/* Material count and, for non-NULL arrays (host or device), a copy of the
 * albedo and emission tables. */
int wf_scene_get_materials(wf_scene scene, int* num_materials, float* albedo_rgb,
                           float* emission_rgb);
/* The 4-wide quantized hierarchy: node count, traversal stack bound and,
 * for a non-NULL host array (num_nodes4 * 20 floats), a copy of the nodes
 * (layout in csrc/scene.h: a header, fp16 child-box planes, references). */
int wf_scene_wide_info(wf_scene scene, int64_t* num_nodes4, int* max_stack4, float* nodes4_host);
/* Scene AABB (getSceneDimensions, base/VulkanglTFModel.cpp:1478-1488). */
int wf_scene_bounds(wf_scene scene, float out_min[3], float out_max[3]);
/* Sizes: triangles, BVH nodes, leaves. */
int wf_scene_info(wf_scene scene, int64_t* num_tris, int64_t* num_nodes,
                  int64_t* num_leaves, int* num_geoms);
/* Copy the built hierarchy to host memory (for inspection / tests):
 * nodes as 16 floats each (child0 lo/hi, child1 lo/hi, child refs as int
 * bit patterns, pad), tri_order[num_nodes + 1] = original triangle of each
 * leaf slot (a split large triangle occupies several slots). */
int wf_scene_download(wf_scene scene, float* nodes_host, int32_t* tri_order_host);

/* ------------------------------------------------------------------- trace */
/* Replaces Tracer::trace (Tracer/Tracer.cpp:185-207) / trace.comp:27-88.
 * For i < n: r = ray_index ? ray_index[i] : i; hits[r] = closest (closest != 0)
 * or any (closest == 0, terminate on first hit) intersection of rays[r] with
 * the scene over [tmin, tmax]; opaque, no face culling.  Closest-hit ties on t
 * resolve to the smaller global triangle index, so results do not depend on
 * traversal order.  tmax < tmin always misses. */
int wf_trace(wf_scene scene, const wf_ray* rays, const uint32_t* ray_index,
             wf_hit* hits, int64_t n, int closest, wf_stream_t stream, float* ms);
/* Tracer::traceSort's reorder + trace (Tracer/Tracer.cpp:80-98 then
 * :185-207, as composed at :209-225 with reorder on) in one pass: the
 * outputs of wf_reorder(idx, in_rays, out_rays, in_s2i, out_s2i, out_i2s)
 * followed by wf_trace(scene, out_rays, NULL, hits, n, closest), bit for bit,
 * with each ray gathered once and traced from registers (configurations
 * without the fused kernel run the two stages). */
int wf_reorder_trace(wf_scene scene, const uint32_t* idx, const wf_ray* in_rays,
                     wf_ray* out_rays, const int32_t* in_s2i, int32_t* out_s2i,
                     int32_t* out_i2s, wf_hit* hits, int64_t n, int closest,
                     wf_stream_t stream, float* ms);

/* Diagnostic variant of wf_trace (same results) that also writes, per ray,
 * the number of BVH nodes visited and triangles tested (u32 arrays indexed
 * like hits).  For BVH quality analysis; not on the rendering path. */
int wf_trace_stats(wf_scene scene, const wf_ray* rays, wf_hit* hits, int64_t n, int closest,
                   uint32_t* node_visits, uint32_t* tri_tests, wf_stream_t stream);

/* -------------------------------------------------------------- ray buffer */
/* Mirror of RayBuffer (Ray/RayBuffer.h:22-70): double-buffered rays and LUTs.
 * All pointers are device pointers owned by the caller (the C++ host class
 * wfpt::RayBuffer manages them).  keys hold key_bytes per element. */
typedef struct wf_ray_buffer {
    wf_ray* rays[2];
    int32_t* index_to_slot[2];
    int32_t* slot_to_index[2];
    wf_hit* results;
    void* keys;
    uint32_t* indices;
    void* sort_scratch;
    int64_t sort_scratch_bytes;
    int64_t size;
    int64_t capacity;
    int32_t swap;        /* RayBuffer::swapBuffers                       */
    int32_t closest_hit; /* RayBuffer::closestHit                         */
} wf_ray_buffer;

/* Tracer::traceSort (Tracer/Tracer.cpp:209-225): keys -> sort -> [reorder,
 * swap] -> trace (indirect through the sorted indices when reorder == 0).
 * times (optional, 4 floats) = {keys, sort, reorder, trace} ms. */
int wf_trace_sort(wf_scene scene, wf_ray_buffer* rb, int key_mode, int digit_bits,
                  const float scene_min[3], const float scene_max[3], int reorder,
                  wf_stream_t stream, float times[4]);

/* ---------------------------------------------------------- ray generation */
/* initSeeds.comp:31-42: seeds[i] = tea(i, frame_index, 16 rounds). */
int wf_init_seeds(uint32_t* seeds, int64_t n, int frame_index, wf_stream_t stream, float* ms);

/* raygenPrimary.comp:44-87 with RayGen::primary (Ray/RayGen.cpp:80-105).
 * screen_to_world: 16 floats, column-major (glm::mat4 layout), the inverse
 * of P*V.  index_to_pixel: PixelTable LUT (n = width*height entries). */
int wf_raygen_primary(const float screen_to_world[16], const float origin[3],
                      const int32_t* index_to_pixel, wf_ray* rays, int width, int height,
                      int sample_index, float max_dist, wf_stream_t stream, float* ms);

/* raygenShadow.comp:68-139 with RayGen::shadow (Ray/RayGen.cpp:107-133).
 * For each input slot j in [batch_begin, batch_end): S rays toward light +
 * radius * (Sobol/Hammersley point + Cranley-Patterson offset).  Seeds are
 * indexed by the path's pixel (in_s2i[j]) so the random numbers follow the
 * path, not its slot. */
int wf_raygen_shadow(const uint32_t* seeds, const wf_ray* in_rays, const wf_hit* in_hits,
                     const int32_t* in_s2i, wf_ray* out_rays, int32_t* out_s2i,
                     int32_t* out_i2s, int64_t batch_begin, int64_t batch_end,
                     int num_samples, const float light[3], float light_radius,
                     wf_stream_t stream, float* ms);

/* raygenPath.comp:78-170 with RayGen::path (Ray/RayGen.cpp:135-186):
 * shade + extend.  Surviving paths are compacted in input order (prefix-sum
 * compaction, deterministic) into out_rays / out_s2i; *d_count (device u32)
 * receives the survivor count and, if host_count != NULL, it is also copied
 * to *host_count (synchronising).  Seeds are indexed by pixel.  decreases
 * (float4 per pixel) *= albedo of the hit material.  With russian_roulette
 * and bounce >= rr_start a path survives with probability
 * q = clamp(max(decrease.rgb), 0.05, 1) and is divided by q. */
int wf_shade_extend(wf_scene scene, uint32_t* seeds, const wf_ray* in_rays,
                    const wf_hit* in_hits, const int32_t* in_s2i, int64_t n,
                    wf_ray* out_rays, int32_t* out_s2i, float* decreases,
                    uint32_t* d_count, int russian_roulette, int bounce, int rr_start,
                    void* scratch, int64_t scratch_bytes, uint32_t* host_count,
                    wf_stream_t stream, float* ms);
int64_t wf_compact_scratch_bytes(int64_t n);

/* ---------------------------------------------------------- reconstruction */
/* Renderer::initDecreases (Renderer/Renderer.cpp:93-104): float4 = 1.0. */
int wf_fill_f32(float* buf, int64_t count, float value, wf_stream_t stream, float* ms);
/* reconstructSmooth.comp:34-97.  pixels[s2i[i]].rgb += decrease * (miss ?
 * background : emission + max(dot(ffN, L), 0.1)) / spp. */
int wf_reconstruct_smooth(wf_scene scene, const wf_ray* rays, const wf_hit* hits,
                          const int32_t* s2i, float* pixels, const float* decreases,
                          int64_t n, const float light[3], const float background[3],
                          int num_samples, wf_stream_t stream, float* ms);
/* reconstructShadow.comp:30-62. */
int wf_reconstruct_shadow(const wf_hit* shadow_hits, const int32_t* in_s2i,
                          const int32_t* shadow_i2s, const float* in_pixels,
                          float* out_pixels, int64_t batch_begin, int64_t batch_end,
                          int num_samples, int replace, wf_stream_t stream, float* ms);
/* interpolateColors.comp:71-100 (progressive average + Reinhard). */
int wf_interpolate_colors(float* frame_pixels, float* pixels, int64_t num_pixels,
                          int frame_index, float key_value, float white_point,
                          wf_stream_t stream, float* ms);
/* countRayHits.comp:22-51: *d_count += number of hits with prim >= 0. */
int wf_count_hits(const wf_hit* hits, int64_t n, uint32_t* d_count, wf_stream_t stream,
                  float* ms);

/* ----------------------------------------------------------------- tuning */
/* Process-wide knobs read by later calls (the autotuner sets them):
 *   "bvh.presplit"     0/1  split references of large triangles (default 1)
 *   "bvh.split_depth"  0-12 max binary splits per triangle (default 8)
 *   "bvh.split_denom"  1-4096 split triangles whose box exceeds the scene
 *                           diagonal / this (default 64; 32-128 measure
 *                           alike on the interior scene, 256+ slower)
 *   "bvh.leaf_max"     1-8  triangles per collapsed leaf (default 1)
 *   "bvh.builder"      0/1  binary hierarchy: 0 radix tree over centroid
 *                           Morton codes (Karras), 1 top-down binned SAH
 *                           (32 bins per axis) (default 1: on the city the
 *                           SAH cost falls from 107 to 72 and the trace
 *                           takes 20% less time, on the interior 117 to 106)
 *   "bvh.collapse"     0/1  4-wide nodes: 0 collapse every second level,
 *                           1 open the largest-area child until 4 children
 *                           (default 1)
 *   "trace.block"      64/128/256  threads per traversal workgroup (default 64)
 *   "sort.onesweep_tiles" >= 0  sorts of more than 8192 pairs in at most this
 *                           many 4096-pair tiles, with digits of at most 8
 *                           bits, run one kernel per pass with decoupled
 *                           look-back (onesweep); the others run
 *                           reduce-then-scan passes (default 64)
 *   "sort.rank_max"    0-1048576  sorts of at most rank_max x passes / 8
 *                           pairs (passes = digit passes the key needs) place
 *                           each pair at its counted rank (keys below it plus
 *                           equal keys before it) in one pass instead of
 *                           digit passes; the order is the same (default
 *                           28000: 60-bit keys in 8-bit digits up to 28,000
 *                           pairs, 30-bit keys up to 14,000; 0 off)
 *   "trace.group64"    >= 0 4-wide traces of at most this many rays run one ray
 *   "trace.group16"         per group of 64 / 16 / 4 lanes: the group splits
 *   "trace.group4"          the top of the hierarchy into subtrees walked in
 *                           parallel with a shared closest distance (latency
 *                           of small wavefronts; defaults 5000 / 24000 / 64000)
 *   "batch.key_layout" 0/1  batched queue sort key below the 6-bit scene prefix:
 *                           0 origin30 << 28 | oct28, 1 the same bits as a 5-D
 *                           Morton code (default 0: on the batched boxes
 *                           scenes, sorted every 4th bounce, both run alike;
 *                           read when a pass is enqueued or captured)
 *   "trace.fuse_reorder" 0/1 wf_reorder_trace as one fused kernel (1, default)
 *                           or as the reorder and trace kernels (0)
 *   "trace.xcd"        >= 0 the 8 XCDs take runs of this many
 *                           consecutive workgroups instead of one each, so
 *                           neighbouring rays share an L2 (default 64: on the
 *                           city's 5.5M-ray bounce the same time with 20%
 *                           fewer HBM bytes; 0 off)
 * Returns WF_ERR_INVALID_ARG for an unknown key or value out of range. */
int wf_set_tuning(const char* key, int value);
int wf_get_tuning(const char* key, int* value);

/* ------------------------------------------------------------- host logic */
/* PixelTable::recalculate (Ray/PixelTable.cpp:21-126): 8x8 Z-order blocks
 * plus stripe tails.  Host arrays of width*height ints. */
int wf_pixel_table(int width, int height, int32_t* index_to_pixel, int32_t* pixel_to_index);

/* Camera (base/camera.hpp:23-45, :108-120) for a first-person camera:
 * writes view, perspective (Y flipped, depth zero-to-one) and
 * screen_to_world = inverse(P*V), each 16 floats column-major. */
int wf_camera_matrices(const float position[3], const float rotation_deg[3], float fov_deg,
                       float aspect, float znear, float zfar, float view[16],
                       float perspective[16], float screen_to_world[16]);

/* -------------------------------------------------------------- renderer */
/* Mirror of Renderer (Renderer/Renderer.h:212-276) driven by an env.json
 * (Environment/AppEnvironment.cpp:10-43 keys and defaults, plus this
 * library's Scene.procedural, Scene.seed, Scene.materialGroups,
 * Renderer.keyMode, Renderer.digitBits, Renderer.reorderPeriod and
 * Renderer.russianRouletteStart). */
typedef struct wf_renderer_s* wf_renderer;
int wf_renderer_create(const char* env_json_path, wf_renderer* out);
int wf_renderer_create_from_string(const char* env_json_text, wf_renderer* out);
int wf_renderer_destroy(wf_renderer r);
/* Builds and binds the scene the env names: Scene.procedural (a generator
 * below) if set, else Scene.filename, a glTF file loaded as by wf_gltf_load
 * (VulkanExample::prepare, WaveFrontPathTracer.cpp:505-535), with its
 * bounds from the file as the reference's model.dimensions.  A relative
 * filename is prefixed with the asset path: "<env file dir>/../assets/" for
 * wf_renderer_create, "./../assets/" (getAssetPath, base/VulkanTools.cpp:13-29)
 * for wf_renderer_create_from_string, or what wf_renderer_set_asset_path set
 * (the reference's resource path).  Scene.materialGroups applies to both. */
int wf_renderer_build_scene(wf_renderer r);
int wf_renderer_set_asset_path(wf_renderer r, const char* dir);
int wf_renderer_set_scene(wf_renderer r, wf_scene scene, const float scene_min[3],
                          const float scene_max[3]);
int wf_renderer_get_scene(wf_renderer r, wf_scene* out);
/* Renders one frame (Renderer::render, Renderer.cpp:552-654). Returns the
 * summed stage GPU ms in *ms_out if non-NULL. */
int wf_renderer_render(wf_renderer r, float* ms_out);
int wf_renderer_reset_frame_index(wf_renderer r);
/* Copies width*height float4 of the accumulated (linear) or tone-mapped
 * framebuffer to host memory. */
int wf_renderer_read_pixels(wf_renderer r, int tone_mapped, float* host_rgba);
/* Image output in place of present.comp (present.comp:19-31) and the
 * swapchain: linear = 0 writes the tone-mapped frame as an 8-bit RGB PNG
 * (UNORM conversion of the rgba8 swapchain image), linear = 1 the
 * accumulated linear estimate as a PFM (float RGB). */
int wf_renderer_save_image(wf_renderer r, const char* path, int linear);
/* The same writers for host float4 images (row 0 = top). */
int wf_write_png(const char* path, const float* host_rgba, int width, int height);
int wf_write_pfm(const char* path, const float* host_rgba, int width, int height);
/* Device pointer of the same framebuffer (width*height float4), valid until
 * the next render or destroy; waits for the renderer's stream. */
int wf_renderer_pixels_device(wf_renderer r, int tone_mapped, const float** out);
/* Path-replay adjoint of the last rendered frame (no counterpart in the
 * reference, whose materials are a constant albedo, raygenPath.comp:138):
 * given dLoss/dpixel (device, width*height float4, rgb used) of the linear
 * frame estimate, writes dLoss/dalbedo and dLoss/demission (device,
 * 3*num_materials floats each, overwritten).  Every path vertex is
 * recomputed under the frame's random numbers; nothing is stored by the
 * forward pass.  replay (device, optional) receives the replayed frame
 * estimate, equal to the rendered one.  Needs Renderer.rayType = path,
 * Russian roulette off, recursionDepth <= 8, and the settings of that frame;
 * frame = the frame index to replay (0 = the last rendered). */
int wf_renderer_adjoint(wf_renderer r, int frame, const float* d_pixels, float* grad_albedo,
                        float* grad_emission, float* replay, float* ms);
int wf_renderer_size(wf_renderer r, int* width, int* height);
/* Stats of the last frame (Renderer.cpp:660-706):
 * counts[3] = primary, shadow, path rays; times[3] = trace ms per type;
 * stage_ms[8] = {keys, sort, reorder, trace_sorted, trace_unsorted,
 *                shade, raygen, reconstruct} summed over the frame for path rays. */
int wf_renderer_stats(wf_renderer r, uint64_t counts[3], float times[3], float stage_ms[8],
                      uint64_t* sorted_rays);
/* Environment lookups (Environment::get*Value). */
/* SortLog (Renderer/Renderer.h:29-46, filled at Renderer.cpp:280-289 and
 * printed per depth by Benchmark.cpp:159-195): per path bounce (kind 0,
 * recursionDepth entries) or per shadow vertex (kind 1, recursionDepth + 1
 * entries, 0 = shadow rays of the primary hits), the last frame's ray count
 * and stage times summed over its passes.  trace_sort_ms is the trace of
 * the sorted wavefront (with the reorder when it is fused into the trace),
 * trace_ms the trace of an unsorted one (the comparison trace that
 * Benchmark.printSortLogs adds, or the only trace when not sorting). */
typedef struct wf_sort_log {
    int64_t ray_count;      /* rays traced                      (rayCount)        */
    int64_t sorted_rays;    /* of which sorted first                              */
    float morton_ms;        /* key generation                   (mortonCodesTime) */
    float sort_ms;          /* radix sort                       (sortTime)        */
    float reorder_ms;       /* reorder, when not fused          (reorderTime)     */
    float trace_sort_ms;    /* trace of the sorted wavefront    (traceSortTime)   */
    float trace_ms;         /* trace of the unsorted wavefront  (traceTime)       */
    float reserved;
} wf_sort_log;
int wf_renderer_sort_logs(wf_renderer r, int kind, wf_sort_log* out, int max_entries,
                          int* num_entries);

/* Decision log (verification, no reference counterpart): when enabled, a
 * frame records per pass p, path vertex v (0 = primary hit, b + 1 = the hit
 * of path bounce b) and pixel the discrete choices its path made there: the
 * global triangle its ray hit (-1 miss, -2 no ray: the path had ended), per
 * shadow sample 1 if the light was visible (0 occluded, 255 no shadow ray),
 * and the sign decisions of the bounce leaving the vertex (bit 0: shading
 * normal flipped to face the ray, bit 1: basis branch of a normal with
 * negative z; 255 no bounce).  wf_renderer_decisions copies them to host
 * arrays (any may be NULL) of spp * (recursionDepth + 1) * width * height
 * int32, that many times numberOfShadowSamples bytes, and that many bytes,
 * indexed [p][v][pixel], [p][v][pixel][sample] and [p][v][pixel]. */
int wf_renderer_set_decision_log(wf_renderer r, int enable);
int wf_renderer_decisions(wf_renderer r, int32_t* hit_tri, uint8_t* shadow_visible,
                          uint8_t* bounce_signs);

int wf_env_get_int(wf_renderer r, const char* name, int* value);
int wf_env_get_float(wf_renderer r, const char* name, float* value);
int wf_env_get_string(wf_renderer r, const char* name, char* buf, int buf_len);
int wf_env_get_vector(wf_renderer r, const char* name, float value[3]);
/* Capture: when enabled, each path bounce's pre-sort wavefront (rays and
 * slot-to-pixel LUT) is copied into renderer-owned device buffers;
 * wf_renderer_captured returns the buffers of slot
 * pass * recursionDepth + bounce (n = 0 where nothing was captured). */
int wf_renderer_set_capture(wf_renderer r, int enable);
int wf_renderer_captured(wf_renderer r, int bounce, const wf_ray** rays,
                         const int32_t** slot_to_index, int64_t* n);

/* ------------------------------------------------------ batched rendering */
/* Many scenes and views of uneven size rendered in one persistent wavefront
 * queue (no counterpart in the reference, which renders one view per
 * Renderer::render, Renderer.cpp:552-654).  Rays of all views share one
 * queue; path rays are sorted by 64-bit keys with the scene index in the
 * top 6 bits (scene | 30-bit origin | 28-bit octahedral direction); counts
 * stay on the device and one pass of the bounce loop is captured in a HIP
 * graph and replayed for every sample pass.  Each view's frame equals the
 * single-view frame loop's (wf_renderer_render, frame 1, Russian roulette
 * off) for the same camera, scene and settings. */
typedef struct {
    int scene;                 /* index into the scenes of wf_batch_create */
    int width, height, spp;
    float position[3], rotation[3];   /* camera (camera.hpp conventions) */
    float fov, znear, zfar;
    float light[3], background[3];
    float shadow_radius;
} wf_view;
typedef struct wf_batch_s* wf_batch;
/* Scenes stay owned by the caller and must outlive the batch. */
int wf_batch_create(const wf_scene* scenes, int num_scenes, const wf_view* views, int num_views,
                    int depth, int shadow_samples, wf_batch* out);
int wf_batch_destroy(wf_batch b);
/* Sort digit width {4, 8, 11}, trace workgroup size {64, 128, 256}, reorder
 * period (sort path rays every k-th bounce; 0 = never). */
int wf_batch_set_tuning(wf_batch b, int digit_bits, int block, int reorder_period);
int wf_batch_get_tuning(wf_batch b, int* digit_bits, int* block, int* reorder_period);
/* Times one full pass (the largest) for every combination of the three
 * knobs and keeps the fastest; *best_ms = its pass time. */
int wf_batch_autotune(wf_batch b, float* best_ms);
/* 1 (default): replay the captured pass graph; 0: launch the same sequence
 * eagerly. */
int wf_batch_set_graph(wf_batch b, int enable);
/* Frame 1 of every view; *ms = GPU time of the frame. */
int wf_batch_render(wf_batch b, float* ms);
/* Linear frame estimate of one view (caller's view index), width*height
 * float4 to host memory. */
int wf_batch_read_pixels(wf_batch b, int view, float* host_rgba);
/* Rays of the last frame {primary, shadow, path} and the queue capacity. */
int wf_batch_stats(wf_batch b, uint64_t counts[3], int64_t* total_pixels);

/* ------------------------------------------------------- procedural scenes */
/* Host-side generators for the five workloads (seeded SplitMix64).  They
 * write host arrays; call first with NULL arrays to get the sizes.
 * Vertex records are {pos.xyz, normal.xyz} (stride 24 bytes).
 * material_groups > 0 regroups the non-emissive triangles into contiguous
 * runs with one material each (material_groups materials in all, emissive
 * materials kept first), as the differentiable workload needs; 0 keeps the
 * generator's materials. */
int wf_procedural_scene(int kind, uint64_t seed, int material_groups, int64_t* num_vertices,
                        int64_t* num_tris,
                        int* num_geoms, int* num_materials, float* vertices,
                        uint32_t* indices, int32_t* geom_first_tri, int32_t* tri_material,
                        float* albedo, float* emission, float scene_min[3],
                        float scene_max[3]);
enum { WF_SCENE_BOXES = 0, WF_SCENE_INTERIOR = 1, WF_SCENE_CITY = 2 };

/* ------------------------------------------------------------ glTF scenes */
/* A glTF 2.0 scene (.gltf with file or base64 buffers, or .glb) as
 * vkglTF::Model::loadFromFile (base/VulkanglTFModel.cpp:767-968,
 * :1178-1338) and the path tracer's acceleration-structure build
 * (WaveFrontPathTracer.cpp:202-275) see it, in the procedural generators'
 * host layout: vertices {pos.xyz, normal.xyz} with positions transformed by
 * their node's world matrix and normals normalised but untransformed (as
 * the shading reads them); u32 indices rebased per primitive (vertexStart);
 * one geometry per indexed primitive in node post-order; tri_material the
 * primitive's glTF material (primitives without one on the extra last
 * material); albedo 0.8 and emission 0 for every material (the reference's
 * constant shading); scene_min/max as Model::getSceneDimensions.  Call
 * first with NULL arrays for the sizes (this reads only the JSON).  Errors:
 * WF_ERR_IO (missing file or buffer), WF_ERR_PARSE (malformed JSON, glb or
 * buffer), WF_ERR_INVALID_ARG (unsupported or inconsistent content). */
int wf_gltf_load(const char* path, int64_t* num_vertices, int64_t* num_tris, int* num_geoms,
                 int* num_materials, float* vertices, uint32_t* indices,
                 int32_t* geom_first_tri, int32_t* tri_material, float* albedo,
                 float* emission, float scene_min[3], float scene_max[3]);

/* The boxes generator with num_boxes boxes (192 triangles each) instead of
 * 64, plus the ground and light quads: the scenes of the batched workload. */
int wf_procedural_boxes(uint64_t seed, int num_boxes, int64_t* num_vertices, int64_t* num_tris,
                        int* num_geoms, int* num_materials, float* vertices, uint32_t* indices,
                        int32_t* geom_first_tri, int32_t* tri_material, float* albedo,
                        float* emission, float scene_min[3], float scene_max[3]);

#ifdef __cplusplus
}
#endif

#endif /* WFPT_H */
