// Status strings, per-thread error messages and the graph-safe fill kernels.
#include "common.h"

#include <cstring>

namespace wfpt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return WF_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_u32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
    return launch_check("fill_u32_kernel");
}

}  // namespace wfpt

using namespace wfpt;

extern "C" const char* wf_status_string(int status) {
    switch (status) {
        case WF_OK: return "ok";
        case WF_ERR_INVALID_ARG: return "invalid argument";
        case WF_ERR_HIP: return "HIP runtime error";
        case WF_ERR_OUT_OF_MEMORY: return "out of device memory";
        case WF_ERR_SCRATCH_TOO_SMALL: return "scratch buffer too small";
        case WF_ERR_TOO_LARGE: return "element count too large";
        case WF_ERR_PARSE: return "parse error";
        case WF_ERR_IO: return "I/O error";
        case WF_ERR_STATE: return "invalid state";
        default: return "unknown status";
    }
}

extern "C" const char* wf_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* wf_version(void) { return "wfpt-mi355x 0.1.0 (gfx950)"; }

extern "C" int wf_fill_f32(float* buf, int64_t count, float value, wf_stream_t stream, float* ms) {
    WF_CHECK_ARG(count >= 0, "wf_fill_f32: count < 0");
    if (count == 0) {
        if (ms) *ms = 0.f;
        return WF_OK;
    }
    WF_CHECK_ARG(buf, "wf_fill_f32: null buffer");
    uint32_t bits;
    std::memcpy(&bits, &value, 4);
    hipStream_t s = as_stream(stream);
    StageTimer t(s, ms);
    WF_TRY(fill_u32(reinterpret_cast<uint32_t*>(buf), count, bits, s));
    return t.finish();
}

namespace wfpt {
Tuning& tuning() {
    static Tuning t;
    return t;
}
static uint64_t g_tuning_generation = 0;
uint64_t tuning_generation() { return g_tuning_generation; }
}  // namespace wfpt

extern "C" int wf_set_tuning(const char* key, int value) {
    WF_CHECK_ARG(key, "wf_set_tuning: null key");
    std::string k(key);
    Tuning& t = tuning();
    if (k == "bvh.presplit" && (value == 0 || value == 1)) t.bvh_presplit = value;
    else if (k == "bvh.split_depth" && value >= 0 && value <= 12) t.bvh_split_depth = value;
    else if (k == "bvh.split_denom" && value >= 1 && value <= 4096) t.bvh_split_denom = value;
    else if (k == "bvh.leaf_max" && value >= 1 && value <= 8) t.bvh_leaf_max = value;
    else if (k == "bvh.collapse" && (value == 0 || value == 1)) t.bvh_collapse = value;
    else if (k == "bvh.builder" && (value == 0 || value == 1)) t.bvh_builder = value;
    else if (k == "trace.block" && (value == 64 || value == 128 || value == 256)) t.trace_block = value;
    else if (k == "trace.xcd" && value >= 0 && value <= 65536) t.trace_xcd = value;
    else if (k == "trace.fuse_reorder" && (value == 0 || value == 1)) t.trace_fuse_reorder = value;
    else if (k == "batch.key_layout" && (value == 0 || value == 1)) t.batch_key_layout = value;
    else if (k == "trace.group64" && value >= 0) t.trace_group64 = value;
    else if (k == "trace.group16" && value >= 0) t.trace_group16 = value;
    else if (k == "trace.group4" && value >= 0) t.trace_group4 = value;
    else if (k == "sort.onesweep_tiles" && value >= 0) t.sort_onesweep_tiles = value;
    else if (k == "sort.rank_max" && value >= 0 && value <= 1048576) t.sort_rank_max = value;
    else return fail(WF_ERR_INVALID_ARG, "wf_set_tuning: unknown key or value out of range: " + k);
    ++g_tuning_generation;
    return WF_OK;
}

extern "C" int wf_get_tuning(const char* key, int* value) {
    WF_CHECK_ARG(key && value, "wf_get_tuning: null argument");
    std::string k(key);
    const Tuning& t = tuning();
    if (k == "bvh.presplit") *value = t.bvh_presplit;
    else if (k == "bvh.split_depth") *value = t.bvh_split_depth;
    else if (k == "bvh.split_denom") *value = t.bvh_split_denom;
    else if (k == "bvh.leaf_max") *value = t.bvh_leaf_max;
    else if (k == "bvh.collapse") *value = t.bvh_collapse;
    else if (k == "bvh.builder") *value = t.bvh_builder;
    else if (k == "trace.block") *value = t.trace_block;
    else if (k == "trace.xcd") *value = t.trace_xcd;
    else if (k == "trace.fuse_reorder") *value = t.trace_fuse_reorder;
    else if (k == "batch.key_layout") *value = t.batch_key_layout;
    else if (k == "trace.group64") *value = t.trace_group64;
    else if (k == "trace.group16") *value = t.trace_group16;
    else if (k == "trace.group4") *value = t.trace_group4;
    else if (k == "sort.onesweep_tiles") *value = t.sort_onesweep_tiles;
    else if (k == "sort.rank_max") *value = t.sort_rank_max;
    else return fail(WF_ERR_INVALID_ARG, "wf_get_tuning: unknown key " + k);
    return WF_OK;
}
