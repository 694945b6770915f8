// C++ host runtime: configuration, camera, pixel order, procedural scenes,
// ray buffers and the frame loop.  Mirrors the reference's host classes
// (Environment, Camera, PixelTable, RayBuffer, PathQueue, Tracer, RayGen,
// Renderer) on top of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/wfpt.h"

namespace wfpt {

// ------------------------------------------------------------ environment
// Environment (Environment/Environment.h:31-71): typed options registered
// with defaults; readEnvFile overrides them from nested JSON objects by the
// dotted option name; unknown JSON keys are ignored.
enum OptType { OPT_INT, OPT_FLOAT, OPT_BOOL, OPT_VECTOR, OPT_STRING };

struct Option {
    OptType type = OPT_STRING;
    std::string name;
    std::vector<std::string> values;
    std::string default_value;
};

class Environment {
  public:
    Environment();  // registers AppEnvironment's options and this library's
    int read_file(const std::string& path);
    int read_string(const std::string& text);
    bool get_int(const std::string& name, int& v) const;
    bool get_float(const std::string& name, float& v) const;
    bool get_bool(const std::string& name, bool& v) const;
    bool get_vector(const std::string& name, float v[3]) const;
    bool get_string(const std::string& name, std::string& v) const;

  private:
    void reg(const std::string& name, const std::string& def, OptType t);
    void reg(const std::string& name, OptType t);
    std::map<std::string, Option> options_;
};

// ------------------------------------------------------------------ camera
struct CameraMatrices {
    float view[16], perspective[16], screen_to_world[16];
};
void camera_matrices(const float pos[3], const float rot_deg[3], float fov_deg, float aspect,
                     float znear, float zfar, CameraMatrices& out);
void pixel_table(int width, int height, std::vector<int32_t>& i2p, std::vector<int32_t>& p2i);

// ------------------------------------------------------- procedural scene
struct HostScene {
    std::vector<float> vertices;       // pos.xyz normal.xyz per vertex
    std::vector<uint32_t> indices;     // 3 per triangle
    std::vector<int32_t> geom_first;   // num_geoms + 1
    std::vector<int32_t> tri_material;
    std::vector<float> albedo, emission;  // 3 per material
    float bmin[3], bmax[3];
};
int procedural_scene(int kind, uint64_t seed, HostScene& out, int num_boxes = 64);
// glTF 2.0 scene (host_gltf.cpp).  out == nullptr: sizes only.
struct GltfSizes {
    int64_t num_vertices = 0, num_tris = 0;
    int num_geoms = 0, num_materials = 0;
};
int load_gltf(const std::string& path, HostScene* out, GltfSizes& sizes);
// Regroups the non-emissive triangles into contiguous runs, one material each.
int regroup_materials(HostScene& s, int groups);

// ----------------------------------------------------------- device memory
struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // vks::util::resizeDiscardBuffer: grow only, contents discarded.
    int ensure(size_t n);
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

}  // namespace wfpt
