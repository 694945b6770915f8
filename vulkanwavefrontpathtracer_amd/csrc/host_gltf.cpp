// glTF 2.0 scenes (Scene.filename) as the path tracer loads them:
// vkglTF::Model::loadFromFile / loadNode (base/VulkanglTFModel.cpp:767-968,
// :1178-1338) feeding the bottom-level acceleration structure
// (projects/WaveFrontPathTracer/WaveFrontPathTracer.cpp:202-275).
//
// What the reference's scene amounts to, and what this loader returns:
// - nodes of the default scene (or scene 0) are visited depth first, each
//   node's children before the node's own mesh (loadNode recurses first and
//   appends the mesh after), so vertices, indices and geometries follow the
//   post-order of linearNodes;
// - every node that carries a mesh appends that mesh's primitives again,
//   each with its own vertex range (vertexStart) and indices rebased onto it;
// - primitives without an index accessor are skipped; primitives with an
//   empty one keep their vertices but make no geometry (indexCount > 0 test,
//   WaveFrontPathTracer.cpp:209, :233);
// - one geometry per remaining primitive with indexCount / 3 triangles; the
//   geometry index is the primitive's position in that order;
// - the acceleration structure applies the node's world matrix
//   (Node::getMatrix = parent chain of T * R * S * matrix, :523-535) to the
//   positions, so positions are returned transformed; normals are returned
//   as the shading reads them: normalised, untransformed
//   (geometrytypes.glsl:33-43, VulkanglTFModel.cpp:885);
// - scene bounds are Model::getSceneDimensions (:1459-1488), which combines
//   the POSITION accessors' min/max as vec4(min, 1) * M, a row vector times
//   the world matrix: that transposes the rotation and drops the
//   translation.  The keys of the sort use these bounds, so they are kept.
//
// Materials: the shading of the reference uses the constant albedo 0.8 and
// no emission (raygenPath.comp:138).  Triangles carry the primitive's glTF
// material index, with primitives without one on an extra last material
// (materials.back(), the default material loadMaterials appends), and every
// material starts at albedo 0.8, emission 0.
//
// Input forms: .gltf with buffers in files next to it or in base64 data
// URIs, and binary .glb.  Accessors honour bufferView.byteStride (the
// reference reads positions and normals as tightly packed floats; for
// tightly packed files both agree).  Unsupported input (sparse accessors,
// non-float positions or normals, non-triangle primitive modes, indices
// outside the primitive's vertices, required extensions) is an error rather
// than undefined behaviour.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "common.h"
#include "host.h"
#include "json.h"

namespace wfpt {

namespace {

constexpr int GL_UNSIGNED_BYTE = 5121, GL_UNSIGNED_SHORT = 5123, GL_UNSIGNED_INT = 5125,
              GL_FLOAT = 5126;
constexpr int MAX_NODE_DEPTH = 1024;

using Mat = std::array<double, 16>;  // column-major, m[col * 4 + row]

Mat identity() {
    Mat m{};
    m[0] = m[5] = m[10] = m[15] = 1.0;
    return m;
}

Mat mul(const Mat& a, const Mat& b) {
    Mat c{};
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + row] * b[col * 4 + k];
            c[col * 4 + row] = s;
        }
    return c;
}

double num(const JValue* v, double def) {
    if (!v || v->kind != JValue::NUM) return def;
    return std::strtod(v->text.c_str(), nullptr);
}
long long inum(const JValue* v, long long def) {
    if (!v || v->kind != JValue::NUM) return def;
    return std::strtoll(v->text.c_str(), nullptr, 10);
}
const JValue* at(const JValue* arr, long long i) {
    if (!arr || arr->kind != JValue::ARR || i < 0 || i >= (long long)arr->arr.size()) return nullptr;
    return &arr->arr[(size_t)i];
}

std::string dir_of(const std::string& path) {
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

std::string percent_decode(const std::string& s) {
    std::string out;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
            std::isxdigit((unsigned char)s[i + 2])) {
            out += (char)std::strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            out += s[i];
        }
    }
    return out;
}

bool base64_decode(const std::string& in, size_t from, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    out.clear();
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < in.size(); ++i) {
        const char c = in[i];
        if (c == '=') break;
        if (std::isspace((unsigned char)c)) continue;
        const int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFFu));
        }
    }
    return true;
}

int read_bytes(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(WF_ERR_IO, "cannot open " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize((size_t)std::max<std::streamoff>(n, 0));
    if (n > 0 && !f.read(reinterpret_cast<char*>(out.data()), n))
        return fail(WF_ERR_IO, "cannot read " + path);
    return WF_OK;
}

uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

class Loader {
  public:
    Loader(const std::string& path, HostScene* out, GltfSizes& sizes)
        : path_(path), dir_(dir_of(path)), out_(out), sz_(sizes) {}

    int run() {
        std::vector<uint8_t> file;
        WF_TRY(read_bytes(path_, file));
        std::string json;
        if (file.size() >= 12 && le32(file.data()) == 0x46546C67u) {  // "glTF": binary container
            WF_TRY(split_glb(file, json));
        } else {
            json.assign(file.begin(), file.end());
        }
        JParser p(json);
        if (!p.parse(root_) || root_.kind != JValue::OBJ)
            return fail(WF_ERR_PARSE, path_ + ": " + p.error());
        if (const JValue* req = root_.find("extensionsRequired"))
            if (req->kind == JValue::ARR && !req->arr.empty())
                return fail(WF_ERR_INVALID_ARG,
                            path_ + ": required extension " + req->arr[0].text + " is not supported");
        nodes_ = root_.find("nodes");
        meshes_ = root_.find("meshes");
        accessors_ = root_.find("accessors");
        views_ = root_.find("bufferViews");
        buffers_ = root_.find("buffers");
        const JValue* materials = root_.find("materials");
        n_gltf_materials_ = materials && materials->kind == JValue::ARR ? (int)materials->arr.size() : 0;

        const JValue* scenes = root_.find("scenes");
        const long long si = inum(root_.find("scene"), 0);
        const JValue* scene = at(scenes, si < 0 ? 0 : si);
        if (!scene) return fail(WF_ERR_INVALID_ARG, path_ + ": no scene to load");
        if (out_) {
            out_->vertices.clear();
            out_->indices.clear();
            out_->tri_material.clear();
            out_->geom_first.assign(1, 0);
        }
        sz_ = GltfSizes{};
        for (int a = 0; a < 3; ++a) {
            bmin_[a] = FLT_MAX;
            bmax_[a] = -FLT_MAX;
        }
        if (const JValue* roots = scene->find("nodes"))
            for (const JValue& n : roots->arr) WF_TRY(load_node(inum(&n, -1), identity(), 0));
        if (sz_.num_tris == 0) return fail(WF_ERR_INVALID_ARG, path_ + ": the scene has no triangles");
        sz_.num_materials = n_gltf_materials_ + 1;
        if (out_) {
            out_->albedo.assign(3 * (size_t)sz_.num_materials, 0.8f);
            out_->emission.assign(3 * (size_t)sz_.num_materials, 0.0f);
            for (int a = 0; a < 3; ++a) {
                out_->bmin[a] = bmin_[a];
                out_->bmax[a] = bmax_[a];
            }
        }
        return WF_OK;
    }

  private:
    struct View {
        const uint8_t* data = nullptr;  // start of the accessor's first element
        size_t stride = 0;
        int64_t count = 0;
    };

    std::string path_, dir_;
    HostScene* out_;
    GltfSizes& sz_;
    JValue root_;
    const JValue *nodes_ = nullptr, *meshes_ = nullptr, *accessors_ = nullptr, *views_ = nullptr,
                 *buffers_ = nullptr;
    std::vector<uint8_t> glb_bin_;
    bool has_glb_bin_ = false;
    std::map<long long, std::vector<uint8_t>> buffer_data_;
    int n_gltf_materials_ = 0;
    float bmin_[3], bmax_[3];

    int split_glb(const std::vector<uint8_t>& f, std::string& json) {
        if (le32(f.data() + 4) != 2) return fail(WF_ERR_PARSE, path_ + ": unsupported glb version");
        const size_t total = std::min<size_t>(le32(f.data() + 8), f.size());
        size_t pos = 12;
        bool have_json = false;
        while (pos + 8 <= total) {
            const size_t len = le32(f.data() + pos);
            const uint32_t type = le32(f.data() + pos + 4);
            pos += 8;
            if (len > total - pos) return fail(WF_ERR_PARSE, path_ + ": truncated glb chunk");
            if (type == 0x4E4F534Au && !have_json) {  // "JSON"
                json.assign(reinterpret_cast<const char*>(f.data() + pos), len);
                have_json = true;
            } else if (type == 0x004E4942u && !has_glb_bin_) {  // "BIN\0"
                glb_bin_.assign(f.begin() + (std::ptrdiff_t)pos, f.begin() + (std::ptrdiff_t)(pos + len));
                has_glb_bin_ = true;
            }
            pos += (len + 3) & ~(size_t)3;
        }
        if (!have_json) return fail(WF_ERR_PARSE, path_ + ": glb without a JSON chunk");
        return WF_OK;
    }

    int buffer(long long bi, const std::vector<uint8_t>** out) {
        auto it = buffer_data_.find(bi);
        if (it != buffer_data_.end()) {
            *out = &it->second;
            return WF_OK;
        }
        const JValue* b = at(buffers_, bi);
        if (!b) return fail(WF_ERR_INVALID_ARG, path_ + ": buffer " + std::to_string(bi) + " missing");
        std::vector<uint8_t> data;
        const JValue* uri = b->find("uri");
        if (!uri) {
            if (bi != 0 || !has_glb_bin_)
                return fail(WF_ERR_INVALID_ARG, path_ + ": buffer without uri or glb data");
            data = glb_bin_;
        } else if (uri->text.compare(0, 5, "data:") == 0) {
            const size_t comma = uri->text.find(',');
            if (comma == std::string::npos ||
                uri->text.substr(0, comma).find(";base64") == std::string::npos ||
                !base64_decode(uri->text, comma + 1, data))
                return fail(WF_ERR_PARSE, path_ + ": unsupported data uri in buffer " + std::to_string(bi));
        } else {
            WF_TRY(read_bytes(dir_ + "/" + percent_decode(uri->text), data));
        }
        const long long declared = inum(b->find("byteLength"), -1);
        if (declared < 0 || (size_t)declared > data.size())
            return fail(WF_ERR_PARSE, path_ + ": buffer " + std::to_string(bi) + " is shorter than byteLength");
        data.resize((size_t)declared);
        *out = &(buffer_data_[bi] = std::move(data));
        return WF_OK;
    }

    // Accessor header checks (counts only) and, when data is wanted, its view.
    int accessor(long long ai, int want_type_size, bool want_data, const char* what, View& v,
                 int* component_type = nullptr) {
        const JValue* a = at(accessors_, ai);
        if (!a) return fail(WF_ERR_INVALID_ARG, path_ + ": " + what + " accessor " + std::to_string(ai) + " missing");
        if (a->find("sparse")) return fail(WF_ERR_INVALID_ARG, path_ + ": sparse accessors are not supported");
        const int ct = (int)inum(a->find("componentType"), 0);
        if (component_type) *component_type = ct;
        v.count = inum(a->find("count"), -1);
        if (v.count < 0) return fail(WF_ERR_PARSE, path_ + ": accessor without count");
        if (!want_data || v.count == 0) return WF_OK;
        const JValue* bv = at(views_, inum(a->find("bufferView"), -1));
        if (!bv) return fail(WF_ERR_INVALID_ARG, path_ + ": accessor without bufferView");
        const std::vector<uint8_t>* buf = nullptr;
        WF_TRY(buffer(inum(bv->find("buffer"), -1), &buf));
        const size_t comp = ct == GL_UNSIGNED_BYTE ? 1 : ct == GL_UNSIGNED_SHORT ? 2 : 4;
        const size_t elem = comp * (size_t)want_type_size;
        const long long view_off = inum(bv->find("byteOffset"), 0);
        const long long view_len = inum(bv->find("byteLength"), -1);
        const long long acc_off = inum(a->find("byteOffset"), 0);
        const long long stride = inum(bv->find("byteStride"), 0);
        if (view_off < 0 || view_len < 0 || acc_off < 0 || stride < 0)
            return fail(WF_ERR_PARSE, path_ + ": " + what + " accessor or view has a negative offset, length or stride");
        v.stride = stride ? (size_t)stride : elem;
        if (v.stride < elem)
            return fail(WF_ERR_PARSE, path_ + ": " + what + " view's byteStride is shorter than an element");
        // Every comparison below is free of overflow: each right-hand side is
        // a difference of sizes already known to be ordered.
        const size_t size = buf->size(), off = (size_t)view_off, len = (size_t)view_len,
                     aoff = (size_t)acc_off;
        const bool fits = off <= size && len <= size - off && aoff <= len && elem <= len - aoff &&
                          (size_t)(v.count - 1) <= (len - aoff - elem) / v.stride;
        if (!fits) return fail(WF_ERR_PARSE, path_ + ": " + what + " accessor runs past its buffer view");
        v.data = buf->data() + off + aoff;
        return WF_OK;
    }

    int load_node(long long ni, const Mat& parent, int depth) {
        const JValue* n = at(nodes_, ni);
        if (!n) return fail(WF_ERR_INVALID_ARG, path_ + ": node " + std::to_string(ni) + " missing");
        if (depth > MAX_NODE_DEPTH) return fail(WF_ERR_INVALID_ARG, path_ + ": node hierarchy has a cycle");
        // local = T * R * S * matrix (Node::localMatrix, VulkanglTFModel.cpp:523-525)
        Mat t = identity(), r = identity(), s = identity(), m = identity();
        if (const JValue* tr = n->find("translation"); tr && tr->arr.size() == 3)
            for (int a = 0; a < 3; ++a) t[12 + a] = num(&tr->arr[a], 0.0);
        if (const JValue* q = n->find("rotation"); q && q->arr.size() == 4) {
            // unit quaternion (x, y, z, w) to its rotation matrix
            const double x = num(&q->arr[0], 0), y = num(&q->arr[1], 0), z = num(&q->arr[2], 0),
                         w = num(&q->arr[3], 1);
            r[0] = 1 - 2 * (y * y + z * z);
            r[1] = 2 * (x * y + w * z);
            r[2] = 2 * (x * z - w * y);
            r[4] = 2 * (x * y - w * z);
            r[5] = 1 - 2 * (x * x + z * z);
            r[6] = 2 * (y * z + w * x);
            r[8] = 2 * (x * z + w * y);
            r[9] = 2 * (y * z - w * x);
            r[10] = 1 - 2 * (x * x + y * y);
        }
        if (const JValue* sc = n->find("scale"); sc && sc->arr.size() == 3)
            for (int a = 0; a < 3; ++a) s[a * 5] = num(&sc->arr[a], 1.0);
        if (const JValue* mm = n->find("matrix"); mm && mm->arr.size() == 16)
            for (int k = 0; k < 16; ++k) m[k] = num(&mm->arr[k], 0.0);
        const Mat world = mul(parent, mul(mul(mul(t, r), s), m));

        if (const JValue* ch = n->find("children"))
            for (const JValue& c : ch->arr) WF_TRY(load_node(inum(&c, -1), world, depth + 1));

        const long long mi = inum(n->find("mesh"), -1);
        if (mi < 0) return WF_OK;
        const JValue* mesh = at(meshes_, mi);
        if (!mesh) return fail(WF_ERR_INVALID_ARG, path_ + ": mesh " + std::to_string(mi) + " missing");
        if (const JValue* prims = mesh->find("primitives"))
            for (const JValue& p : prims->arr) WF_TRY(load_primitive(p, world));
        return WF_OK;
    }

    int load_primitive(const JValue& p, const Mat& world) {
        const long long ii = inum(p.find("indices"), -1);
        if (ii < 0) return WF_OK;  // VulkanglTFModel.cpp:813-815
        const long long mode = inum(p.find("mode"), 4);
        if (mode != 4)
            return fail(WF_ERR_INVALID_ARG, path_ + ": primitive mode " + std::to_string(mode) +
                                                " is not a triangle list");
        const JValue* attrs = p.find("attributes");
        const JValue* pos_a = attrs ? attrs->find("POSITION") : nullptr;
        if (!pos_a) return fail(WF_ERR_INVALID_ARG, path_ + ": primitive without POSITION");
        const JValue* nrm_a = attrs->find("NORMAL");
        const bool fill = out_ != nullptr;

        View pv, nv, iv;
        int ct = 0;
        WF_TRY(accessor(inum(pos_a, -1), 3, fill, "POSITION", pv, &ct));
        if (ct != GL_FLOAT) return fail(WF_ERR_INVALID_ARG, path_ + ": POSITION must be float");
        if (nrm_a) {
            WF_TRY(accessor(inum(nrm_a, -1), 3, fill, "NORMAL", nv, &ct));
            if (ct != GL_FLOAT) return fail(WF_ERR_INVALID_ARG, path_ + ": NORMAL must be float");
            if (nv.count < pv.count) return fail(WF_ERR_INVALID_ARG, path_ + ": fewer normals than positions");
        }
        WF_TRY(accessor(ii, 1, fill, "index", iv, &ct));
        if (ct != GL_UNSIGNED_BYTE && ct != GL_UNSIGNED_SHORT && ct != GL_UNSIGNED_INT)
            return fail(WF_ERR_INVALID_ARG, path_ + ": index component type " + std::to_string(ct) +
                                                " not supported");

        const int64_t vertex_start = sz_.num_vertices;
        const int64_t ntri = iv.count / 3;
        if (vertex_start + pv.count > (int64_t)UINT32_MAX)
            return fail(WF_ERR_TOO_LARGE, path_ + ": more than 2^32 vertices");

        // bounds contribution (Model::getNodeDimensions): vec4(min, 1) * M,
        // from the accessor's min/max, or from the data where they are absent
        if (fill && pv.count > 0) {
            const JValue* pa = at(accessors_, inum(pos_a, -1));
            const JValue* amin = pa->find("min");
            const JValue* amax = pa->find("max");
            float lo[3], hi[3];
            if (amin && amax && amin->arr.size() == 3 && amax->arr.size() == 3) {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = (float)num(&amin->arr[a], 0.0);
                    hi[a] = (float)num(&amax->arr[a], 0.0);
                }
            } else {
                for (int a = 0; a < 3; ++a) {
                    lo[a] = FLT_MAX;
                    hi[a] = -FLT_MAX;
                }
                for (int64_t v = 0; v < pv.count; ++v) {
                    float q[3];
                    std::memcpy(q, pv.data + (size_t)v * pv.stride, sizeof q);
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], q[a]);
                        hi[a] = std::max(hi[a], q[a]);
                    }
                }
            }
            for (int j = 0; j < 3; ++j) {
                double l = world[j * 4 + 3], h = world[j * 4 + 3];
                for (int i = 0; i < 3; ++i) {
                    l += (double)lo[i] * world[j * 4 + i];
                    h += (double)hi[i] * world[j * 4 + i];
                }
                bmin_[j] = std::min(bmin_[j], (float)l);
                bmax_[j] = std::max(bmax_[j], (float)h);
            }
        }

        if (fill) {
            auto& V = out_->vertices;
            V.reserve(V.size() + 6 * (size_t)pv.count);
            for (int64_t v = 0; v < pv.count; ++v) {
                float q[3], nr[3] = {0.0f, 0.0f, 0.0f};
                std::memcpy(q, pv.data + (size_t)v * pv.stride, sizeof q);
                if (nrm_a) std::memcpy(nr, nv.data + (size_t)v * nv.stride, sizeof nr);
                for (int r = 0; r < 3; ++r) {
                    const double w = world[r] * q[0] + world[4 + r] * q[1] + world[8 + r] * q[2] + world[12 + r];
                    V.push_back((float)w);
                }
                // glm::normalize: v * inversesqrt(dot(v, v)); a missing
                // normal stays zero
                const float d = nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2];
                const float is = d > 0.0f ? 1.0f / std::sqrt(d) : 0.0f;
                for (int a = 0; a < 3; ++a) V.push_back(nr[a] * is);
            }
            if (ntri > 0) {
                auto& I = out_->indices;
                for (int64_t k = 0; k < 3 * ntri; ++k) {
                    const uint8_t* e = iv.data + (size_t)k * iv.stride;
                    uint32_t x = ct == GL_UNSIGNED_BYTE    ? e[0]
                                 : ct == GL_UNSIGNED_SHORT ? (uint32_t)(e[0] | (e[1] << 8))
                                                           : le32(e);
                    if ((int64_t)x >= pv.count)
                        return fail(WF_ERR_INVALID_ARG, path_ + ": index " + std::to_string(x) +
                                                            " outside its primitive's vertices");
                    I.push_back(x + (uint32_t)vertex_start);
                }
                const long long mat = inum(p.find("material"), -1);
                const int32_t m = mat >= 0 && mat < n_gltf_materials_ ? (int32_t)mat : n_gltf_materials_;
                out_->tri_material.insert(out_->tri_material.end(), (size_t)ntri, m);
            }
        }
        sz_.num_vertices += pv.count;
        if (ntri > 0) {
            sz_.num_tris += ntri;
            ++sz_.num_geoms;
            if (sz_.num_tris > (int64_t)INT32_MAX)
                return fail(WF_ERR_TOO_LARGE, path_ + ": more than 2^31 triangles");
            if (fill) out_->geom_first.push_back((int32_t)sz_.num_tris);
        }
        return WF_OK;
    }
};

}  // namespace

int load_gltf(const std::string& path, HostScene* out, GltfSizes& sizes) {
    Loader l(path, out, sizes);
    return l.run();
}

}  // namespace wfpt

using namespace wfpt;

extern "C" int wf_gltf_load(const char* path, int64_t* num_vertices, int64_t* num_tris,
                            int* num_geoms, int* num_materials, float* vertices,
                            uint32_t* indices, int32_t* geom_first_tri, int32_t* tri_material,
                            float* albedo, float* emission, float scene_min[3],
                            float scene_max[3]) {
    WF_CHECK_ARG(path, "wf_gltf_load: null path");
    const bool want = vertices || indices || geom_first_tri || tri_material || albedo || emission ||
                      scene_min || scene_max;
    GltfSizes sz;
    HostScene hs;
    WF_TRY(load_gltf(path, want ? &hs : nullptr, sz));
    if (num_vertices) *num_vertices = sz.num_vertices;
    if (num_tris) *num_tris = sz.num_tris;
    if (num_geoms) *num_geoms = sz.num_geoms;
    if (num_materials) *num_materials = sz.num_materials;
    if (!want) return WF_OK;
    if (vertices) std::memcpy(vertices, hs.vertices.data(), sizeof(float) * hs.vertices.size());
    if (indices) std::memcpy(indices, hs.indices.data(), sizeof(uint32_t) * hs.indices.size());
    if (geom_first_tri)
        std::memcpy(geom_first_tri, hs.geom_first.data(), sizeof(int32_t) * hs.geom_first.size());
    if (tri_material)
        std::memcpy(tri_material, hs.tri_material.data(), sizeof(int32_t) * hs.tri_material.size());
    if (albedo) std::memcpy(albedo, hs.albedo.data(), sizeof(float) * hs.albedo.size());
    if (emission) std::memcpy(emission, hs.emission.data(), sizeof(float) * hs.emission.size());
    if (scene_min) std::memcpy(scene_min, hs.bmin, sizeof hs.bmin);
    if (scene_max) std::memcpy(scene_max, hs.bmax, sizeof hs.bmax);
    return WF_OK;
}
