// mini_json.hpp - a small JSON reader for the configuration tests (objects,
// arrays, strings with the common escapes, numbers, true / false / null).  Test
// infrastructure only: it fills the plain config structs of
// include/icarion_config.hpp from the reference's example files, as the
// reference's own jsoncpp loaders do.
#pragma once

#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mini_json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool has(const std::string& k) const { return kind == Object && obj.count(k) > 0; }
    const Value& operator[](const std::string& k) const {
        static const Value null_value;
        if (kind != Object) return null_value;
        const auto it = obj.find(k);
        return it == obj.end() ? null_value : it->second;
    }
    const Value& operator[](size_t i) const { return arr.at(i); }
    size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
    bool is_number() const { return kind == Number; }
    double as_double(double dflt = 0.0) const { return kind == Number ? num : dflt; }
    std::string as_string(const std::string& dflt = "") const { return kind == String ? str : dflt; }
    bool as_bool(bool dflt = false) const { return kind == Bool ? b : dflt; }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& what) {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(i_) + ": " + what);
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
    }
    bool lit(const char* w) {
        const std::string t(w);
        if (s_.compare(i_, t.size(), t) == 0) {
            i_ += t.size();
            return true;
        }
        return false;
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        const char c = s_[i_];
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            ws();
            if (s_[i_] == '}') { ++i_; return v; }
            for (;;) {
                ws();
                if (s_[i_] != '"') fail("expected a key");
                const std::string k = string();
                ws();
                if (s_[i_] != ':') fail("expected ':'");
                ++i_;
                v.obj[k] = value();
                ws();
                if (s_[i_] == ',') { ++i_; continue; }
                if (s_[i_] == '}') { ++i_; return v; }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            ws();
            if (s_[i_] == ']') { ++i_; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (s_[i_] == ',') { ++i_; continue; }
                if (s_[i_] == ']') { ++i_; return v; }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.str = string();
            return v;
        }
        if (lit("true")) { v.kind = Value::Bool; v.b = true; return v; }
        if (lit("false")) { v.kind = Value::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        char* end = nullptr;
        v.num = std::strtod(s_.c_str() + i_, &end);
        if (end == s_.c_str() + i_) fail("bad value");
        i_ = static_cast<size_t>(end - s_.c_str());
        v.kind = Value::Number;
        return v;
    }
    std::string string() {
        ++i_;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                const char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        // keep the code point's UTF-8 bytes (BMP only)
                        const unsigned cp = static_cast<unsigned>(std::stoul(s_.substr(i_, 4), nullptr, 16));
                        i_ += 4;
                        if (cp < 0x80) out += static_cast<char>(cp);
                        else if (cp < 0x800) {
                            out += static_cast<char>(0xC0 | (cp >> 6));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        } else {
                            out += static_cast<char>(0xE0 | (cp >> 12));
                            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;
        return out;
    }
};

inline Value parse_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    return Parser(text).parse();
}

}  // namespace mini_json
