// Internal helpers shared by the HIP translation units of libwfpt.
//
// All device code is compiled with -ffp-contract=off: every multiply-add that
// should fuse is written as an explicit fmaf(), so the CPU restatement can
// reproduce the arithmetic bit for bit.  fp32 division and sqrt are correctly
// rounded (clang's HIP default).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/wfpt.h"

namespace wfpt {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define WF_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return ::wfpt::fail(WF_ERR_HIP, std::string(#call) + ": " +              \
                                                hipGetErrorString(e_));              \
    } while (0)

#define WF_CHECK_ARG(cond, msg)                                                      \
    do {                                                                             \
        if (!(cond)) return ::wfpt::fail(WF_ERR_INVALID_ARG, msg);                   \
    } while (0)

#define WF_TRY(expr)                                                                 \
    do {                                                                             \
        int s_ = (expr);                                                             \
        if (s_ != WF_OK) return s_;                                                  \
    } while (0)

inline hipStream_t as_stream(wf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Brackets the launches of one entry point with hipEvents when the caller
// asked for a time (ms != nullptr), like ComputePass::launchTimed.
class StageTimer {
  public:
    StageTimer(hipStream_t s, float* ms) : s_(s), ms_(ms) {
        if (ms_) {
            if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) {
                ms_ = nullptr;
                return;
            }
            (void)hipEventRecord(a_, s_);
        }
    }
    int finish() {
        if (!ms_) return WF_OK;
        WF_HIP(hipEventRecord(b_, s_));
        WF_HIP(hipEventSynchronize(b_));
        float t = 0.f;
        WF_HIP(hipEventElapsedTime(&t, a_, b_));
        *ms_ = t;
        return WF_OK;
    }
    ~StageTimer() {
        if (a_) (void)hipEventDestroy(a_);
        if (b_) (void)hipEventDestroy(b_);
    }

  private:
    hipStream_t s_;
    float* ms_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

inline int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(WF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return WF_OK;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------- device helpers
struct float3_ {
    float x, y, z;
};

__device__ __forceinline__ float4 ld_f4(const float* p) {
    return *reinterpret_cast<const float4*>(p);
}

// Load one 32-byte Ray record as two 16-byte loads.
__device__ __forceinline__ void load_ray(const wf_ray* rays, int64_t i, float4& a, float4& b) {
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    a = p[0];
    b = p[1];
}
__device__ __forceinline__ void store_ray(wf_ray* rays, int64_t i, float4 a, float4 b) {
    float4* p = reinterpret_cast<float4*>(rays + i);
    p[0] = a;
    p[1] = b;
}

// Wave64 primitives.
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }
// Workgroup b of a g-workgroup grid renumbered so that the 8 XCDs, which
// receive workgroups round-robin (b mod 8), each take runs of k consecutive
// workgroups: run c of XCD x is the grid's run 8c + x.  Consecutive
// workgroups of a sorted wavefront then share an L2 while every XCD still
// samples the whole wavefront.  Workgroups past the last whole round of
// 8k keep their index.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t g, int64_t k) {
    constexpr int64_t NX = 8;
    const int64_t whole = (g / (NX * k)) * (NX * k);
    if (b >= whole) return b;
    const int64_t x = b % NX, p = b / NX;
    return ((p / k) * NX + x) * k + (p % k);
}

// Number of set bits of `mask` below the calling lane.
__device__ __forceinline__ unsigned popc_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Look-back status words: 2 flag bits + 30-bit count.
constexpr uint32_t LB_FLAG_AGG = 1u << 30;
constexpr uint32_t LB_FLAG_INC = 2u << 30;
constexpr uint32_t LB_MASK = (1u << 30) - 1u;

__device__ __forceinline__ uint32_t lb_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fill kernel used instead of hipMemsetAsync so that the clears stay valid
// inside captured graphs.
__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v);
int fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// Decision log of the frame loop (reconstruct.hip): out[s2i[i]] = global
// triangle hit by ray i (-1 miss); out[pixel * S + sample] = 1 where the
// shadow ray reached the light.
int log_hits(const wf_hit* hits, const int32_t* s2i, int64_t n, const int32_t* geom_first,
             int32_t* out, hipStream_t s);
int log_shadow(const wf_hit* shadow_hits, const int32_t* shadow_i2s, const int32_t* in_s2i,
               int64_t begin, int64_t end, int S, uint8_t* out, hipStream_t s);
// wf_shade_extend that also writes, per shaded pixel, its bounce's sign
// decisions (bit 0: normal flipped to face the ray; bit 1: basis branch of a
// negative z) to sign_log (nullptr: none).
int shade_extend_logged(wf_scene scene, uint32_t* seeds, const wf_ray* in_rays,
                        const wf_hit* in_hits, const int32_t* in_s2i, int64_t n, wf_ray* out_rays,
                        int32_t* out_s2i, float* decreases, uint32_t* d_count,
                        int russian_roulette, int bounce, int rr_start, void* scratch,
                        int64_t scratch_bytes, uint32_t* host_count, wf_stream_t stream, float* ms,
                        uint8_t* sign_log);

}  // namespace wfpt

namespace wfpt {
// Process-wide tuning knobs (wf_set_tuning).
struct Tuning {
    int bvh_presplit = 1;
    int bvh_split_depth = 8;
    int bvh_split_denom = 64;   // split triangles whose box exceeds diagonal / this
    int bvh_leaf_max = 1;
    int bvh_collapse = 1;       // 4-wide nodes: 0 every second level; 1 greedy largest-area
    int bvh_builder = 1;        // binary hierarchy: 0 radix tree (Karras), 1 binned SAH
    int trace_block = 64;
    int sort_onesweep_tiles = 64;  // sorts of at most this many 4096-element tiles use onesweep
    int sort_rank_max = 28000;  // sorts of at most this many pairs x passes / 8 count each rank directly (0 off)
    int trace_group64 = 5000;   // wavefronts of at most this many rays: 64 lanes per ray
    int trace_group16 = 24000;  // ... 16 lanes per ray
    int trace_group4 = 64000;   // ... 4 lanes per ray
    int batch_key_layout = 0;   // batched queue key below the scene prefix: 0 concatenated, 1 5-D Morton
    int trace_fuse_reorder = 1;  // wf_reorder_trace: one fused kernel where available
    int trace_xcd = 64;         // runs of this many consecutive workgroups per XCD (0 off)
};
Tuning& tuning();
// Incremented by every successful wf_set_tuning: a captured graph records the
// knobs read at capture time, so its owner recaptures when this changes.
uint64_t tuning_generation();
}  // namespace wfpt
