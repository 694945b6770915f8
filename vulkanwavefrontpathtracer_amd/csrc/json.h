// Minimal JSON reader shared by the env.json registry and the glTF loader:
// objects keep their keys in file order, numbers keep their text, strings
// are unescaped (\uXXXX below 0x80 only, which is all either format uses
// for keys and values that matter here).
#pragma once

#include <cctype>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace wfpt {

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
    std::string error() const {
        return depth_ > MAX_DEPTH ? "JSON nesting deeper than " + std::to_string(MAX_DEPTH) + " levels"
                                  : "JSON parse error at offset " + std::to_string(i_);
    }

    // Arrays and objects nest at most this deep (the parser recurses once
    // per level; deeper input is rejected instead of exhausting the stack).
    static constexpr int MAX_DEPTH = 256;

  private:
    const std::string& s_;
    size_t i_ = 0;
    int depth_ = 0;
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
        if (c == '{' || c == '[') {
            if (++depth_ > MAX_DEPTH) return false;
            const bool ok = c == '{' ? object(v) : array(v);
            --depth_;
            return ok;
        }
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

}  // namespace wfpt
