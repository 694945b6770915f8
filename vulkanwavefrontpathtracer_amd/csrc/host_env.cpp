// Configuration registry for env.json files (JSON reader in json.h).
//
// Behaviour follows Environment.cpp:20-69 (value filtering of defaults),
// :121-197 (getters: first value, else default) and :286-377 (readEnvFile:
// nested lookup by dotted name; strings kept verbatim, booleans as "1"/"0",
// other scalars as their JSON text, arrays element-wise).
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "common.h"
#include "host.h"
#include "json.h"

namespace wfpt {

namespace {

// Environment.cpp:20-69.
bool filter_value(const std::string& value, std::string& out, OptType type) {
    if (type == OPT_INT) {
        try {
            out = std::to_string(std::stoi(value));
        } catch (...) {
            return false;
        }
    } else if (type == OPT_FLOAT) {
        try {
            out = std::to_string(std::stof(value));
        } catch (...) {
            return false;
        }
    } else if (type == OPT_BOOL) {
        std::string v = value;
        for (auto& ch : v) ch = (char)std::tolower((unsigned char)ch);
        if (v == "true" || v == "yes" || v == "on" || v == "1") out = "1";
        else if (v == "false" || v == "no" || v == "off" || v == "0") out = "0";
        else return false;
    } else if (type == OPT_VECTOR) {
        std::stringstream ss(value);
        float f[3];
        int count = 0;
        while (count < 3 && (ss >> f[count])) ++count;
        if (count != 3) return false;
        out = std::to_string(f[0]) + " " + std::to_string(f[1]) + " " + std::to_string(f[2]);
    } else {
        out = value;
        if (value.empty()) return false;
    }
    return true;
}

}  // namespace

void Environment::reg(const std::string& name, const std::string& def, OptType t) {
    Option o;
    o.name = name;
    o.type = t;
    if (!filter_value(def, o.default_value, t)) o.default_value.clear();
    options_[name] = o;
}
void Environment::reg(const std::string& name, OptType t) {
    Option o;
    o.name = name;
    o.type = t;
    options_[name] = o;
}

Environment::Environment() {
    // AppEnvironment::registerOptions (Environment/AppEnvironment.cpp:10-43).
    reg("Application.mode", "interactive", OPT_STRING);
    reg("Benchmark.printSortLogs", "0", OPT_BOOL);
    reg("Resolution.width", "1024", OPT_INT);
    reg("Resolution.height", "768", OPT_INT);
    reg("Scene.filename", OPT_STRING);
    reg("Scene.light", "5.0 -2.0 1.0", OPT_VECTOR);
    reg("Scene.backgroundcolor", "0.0 0.0 1.0", OPT_VECTOR);
    reg("Scene.headlight", "0", OPT_BOOL);
    reg("Renderer.numberOfPrimarySamples", "1", OPT_INT);
    reg("Renderer.numberOfShadowSamples", "2", OPT_INT);
    reg("Renderer.shadowRadius", "0.001", OPT_FLOAT);
    reg("Renderer.rayType", "primary", OPT_STRING);
    reg("Renderer.recursionDepth", "1", OPT_INT);
    reg("Renderer.keyValue", "0.6", OPT_FLOAT);
    reg("Renderer.whitePoint", "2.0", OPT_FLOAT);
    reg("Renderer.russianRoulette", "0", OPT_BOOL);
    reg("Renderer.sortShadowRays", "0", OPT_BOOL);
    reg("Renderer.reorderShadowRays", "0", OPT_BOOL);
    reg("Renderer.sortPathRays", "0", OPT_BOOL);
    reg("Renderer.reorderPathRays", "0", OPT_BOOL);
    reg("Camera.position", "0.0 0.0 0.0", OPT_VECTOR);
    reg("Camera.rotation", "0.0 0.0 0.0", OPT_VECTOR);
    reg("Camera.nearPlane", "0.001", OPT_FLOAT);
    reg("Camera.farPlane", "3.0", OPT_FLOAT);
    reg("Camera.fieldOfView", "45.0", OPT_FLOAT);
    // This library's keys (procedural workloads and sort settings).
    reg("Scene.procedural", OPT_STRING);          // boxes | interior | city
    reg("Scene.seed", "1000", OPT_INT);
    reg("Scene.materialGroups", "0", OPT_INT);     // > 0: contiguous material groups
    reg("Scene.splitDepth", "-1", OPT_INT);        // >= 0: bvh.split_depth for this scene's build
    reg("Renderer.keyMode", "aila32", OPT_STRING);  // aila32 | origin30 | origin30_oct30 | origin_dir_morton60 | origin_dir_morton30
    reg("Renderer.digitBits", "8", OPT_INT);        // 4 | 8 | 11
    reg("Renderer.reorderPeriod", "1", OPT_INT);    // sort every k-th bounce
    reg("Renderer.russianRouletteStart", "3", OPT_INT);
}

int Environment::read_string(const std::string& text) {
    JValue root;
    JParser p(text);
    if (!p.parse(root)) return fail(WF_ERR_PARSE, "env.json: " + p.error());
    for (auto& it : options_) {
        Option& opt = it.second;
        const JValue* cur = &root;
        std::stringstream ss(opt.name);
        std::string seg;
        bool found = true;
        while (std::getline(ss, seg, '.')) {
            if (cur->kind != JValue::OBJ || !(cur = cur->find(seg))) {
                found = false;
                break;
            }
        }
        if (!found) continue;
        opt.values.clear();
        auto push = [&](const JValue& el) {
            if (el.kind == JValue::STR) opt.values.push_back(el.text);
            else if (el.kind == JValue::BOOL) opt.values.push_back(el.b ? "1" : "0");
            else if (el.kind == JValue::NUM) opt.values.push_back(el.text);
            else opt.values.push_back("");
        };
        if (cur->kind == JValue::ARR)
            for (auto& el : cur->arr) push(el);
        else
            push(*cur);
    }
    return WF_OK;
}

int Environment::read_file(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(WF_ERR_IO, "env.json: cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return read_string(ss.str());
}

namespace {
const std::string* first_value(const std::map<std::string, Option>& m, const std::string& n) {
    auto it = m.find(n);
    if (it == m.end()) return nullptr;
    const Option& o = it->second;
    if (!o.values.empty() && !o.values.front().empty()) return &o.values.front();
    if (!o.default_value.empty()) return &o.default_value;
    return nullptr;
}
}  // namespace

bool Environment::get_int(const std::string& name, int& v) const {
    const std::string* s = first_value(options_, name);
    if (!s) return false;
    try {
        v = std::stoi(*s);
    } catch (...) {
        return false;
    }
    return true;
}
bool Environment::get_float(const std::string& name, float& v) const {
    const std::string* s = first_value(options_, name);
    if (!s) return false;
    try {
        v = std::stof(*s);
    } catch (...) {
        return false;
    }
    return true;
}
bool Environment::get_bool(const std::string& name, bool& v) const {
    const std::string* s = first_value(options_, name);
    if (!s) return false;
    std::string f;
    if (filter_value(*s, f, OPT_BOOL)) {
        v = f == "1";
        return true;
    }
    try {
        v = std::stoi(*s) != 0;
    } catch (...) {
        return false;
    }
    return true;
}
bool Environment::get_vector(const std::string& name, float v[3]) const {
    const std::string* s = first_value(options_, name);
    if (!s) return false;
    std::stringstream ss(*s);
    ss >> v[0] >> v[1] >> v[2];
    return !ss.fail();
}
bool Environment::get_string(const std::string& name, std::string& v) const {
    const std::string* s = first_value(options_, name);
    if (!s) return false;
    v = *s;
    return true;
}

int DeviceBuffer::ensure(size_t n) {
    if (n <= bytes && p) return WF_OK;
    release();
    size_t alloc = n ? n : 16;
    hipError_t e = hipMalloc(&p, alloc);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(WF_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = alloc;
    return WF_OK;
}

}  // namespace wfpt
