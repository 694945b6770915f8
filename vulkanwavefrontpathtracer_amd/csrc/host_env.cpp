// Configuration registry and a small JSON reader for env.json files.
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

namespace wfpt {

namespace {

struct JValue {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    std::string text;  // number text or string contents
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;
    const JValue* find(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JParser {
  public:
    explicit JParser(const std::string& s) : s_(s) {}
    bool parse(JValue& v) {
        ws();
        if (!value(v)) return false;
        ws();
        return i_ == s_.size();
    }
    std::string error() const { return "JSON parse error at offset " + std::to_string(i_); }

  private:
    const std::string& s_;
    size_t i_ = 0;
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(i_, n, w) != 0) return false;
        i_ += n;
        return true;
    }
    bool value(JValue& v) {
        if (i_ >= s_.size()) return false;
        char c = s_[i_];
        if (c == '{') return object(v);
        if (c == '[') return array(v);
        if (c == '"') {
            v.kind = JValue::STR;
            return string(v.text);
        }
        if (lit("true")) {
            v.kind = JValue::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JValue::BOOL;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.kind = JValue::NUL;
            return true;
        }
        return number(v);
    }
    bool number(JValue& v) {
        size_t st = i_;
        if (i_ < s_.size() && (s_[i_] == '-' || s_[i_] == '+')) ++i_;
        bool digits = false;
        while (i_ < s_.size() && (std::isdigit((unsigned char)s_[i_]) || s_[i_] == '.' ||
                                  s_[i_] == 'e' || s_[i_] == 'E' || s_[i_] == '-' || s_[i_] == '+')) {
            digits = digits || std::isdigit((unsigned char)s_[i_]);
            ++i_;
        }
        if (!digits) {
            i_ = st;
            return false;
        }
        v.kind = JValue::NUM;
        v.text = s_.substr(st, i_ - st);
        return true;
    }
    bool string(std::string& out) {
        ++i_;  // opening quote
        out.clear();
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) return false;
                char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        if (i_ + 4 > s_.size()) return false;
                        out += (char)std::strtol(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) return false;
        ++i_;
        return true;
    }
    bool array(JValue& v) {
        v.kind = JValue::ARR;
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == ']') {
            ++i_;
            return true;
        }
        while (true) {
            JValue e;
            ws();
            if (!value(e)) return false;
            v.arr.push_back(std::move(e));
            ws();
            if (i_ < s_.size() && s_[i_] == ',') {
                ++i_;
                continue;
            }
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return true;
            }
            return false;
        }
    }
    bool object(JValue& v) {
        v.kind = JValue::OBJ;
        ++i_;
        ws();
        if (i_ < s_.size() && s_[i_] == '}') {
            ++i_;
            return true;
        }
        while (true) {
            ws();
            if (i_ >= s_.size() || s_[i_] != '"') return false;
            std::string key;
            if (!string(key)) return false;
            ws();
            if (i_ >= s_.size() || s_[i_] != ':') return false;
            ++i_;
            ws();
            JValue e;
            if (!value(e)) return false;
            v.obj.emplace_back(std::move(key), std::move(e));
            ws();
            if (i_ < s_.size() && s_[i_] == ',') {
                ++i_;
                continue;
            }
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return true;
            }
            return false;
        }
    }
};

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
    reg("Renderer.keyMode", "aila32", OPT_STRING);  // aila32 | origin30 | origin30_oct30
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
