// json_reader.hpp -- a small JSON reader (objects, arrays, strings, numbers,
// literals) for the basis-set and dispatch-table files of the host layer.
#pragma once

#include <libaccint/core.hpp>

#include <cctype>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace libaccint::detail {

struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;
    const JValue* get(const std::string& k) const {
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JParser {
public:
    JParser(const std::string& s, const char* what) : s_(s), what_(what) {}
    JValue parse() {
        JValue v = value();
        ws();
        if (p_ != s_.size()) err("trailing characters");
        return v;
    }

private:
    [[noreturn]] void err(const char* m) {
        throw InvalidArgumentException(std::string("Failed to parse ") + what_ + ": " + m + " at offset " +
                                       std::to_string(p_));
    }
    void ws() {
        while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
    }
    char peek() {
        ws();
        if (p_ >= s_.size()) err("unexpected end");
        return s_[p_];
    }
    void expect(char c) {
        if (peek() != c) err("unexpected character");
        ++p_;
    }
    std::string string_lit() {
        expect('"');
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                if (p_ >= s_.size()) err("bad escape");
                char e = s_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (p_ + 4 > s_.size()) err("bad unicode escape");
                        const unsigned cp = std::stoul(s_.substr(p_, 4), nullptr, 16);
                        p_ += 4;
                        if (cp < 0x80) out += char(cp);
                        else if (cp < 0x800) {
                            out += char(0xC0 | (cp >> 6));
                            out += char(0x80 | (cp & 0x3F));
                        } else {
                            out += char(0xE0 | (cp >> 12));
                            out += char(0x80 | ((cp >> 6) & 0x3F));
                            out += char(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) err("unterminated string");
        ++p_;
        return out;
    }
    JValue value() {
        const char c = peek();
        JValue v;
        if (c == '{') {
            ++p_;
            v.kind = JValue::Object;
            if (peek() == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                std::string k = string_lit();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
                const char d = peek();
                ++p_;
                if (d == '}') break;
                if (d != ',') err("expected , or }");
            }
        } else if (c == '[') {
            ++p_;
            v.kind = JValue::Array;
            if (peek() == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                const char d = peek();
                ++p_;
                if (d == ']') break;
                if (d != ',') err("expected , or ]");
            }
        } else if (c == '"') {
            v.kind = JValue::String;
            v.str = string_lit();
        } else if (s_.compare(p_, 4, "true") == 0) {
            v.kind = JValue::Bool;
            v.b = true;
            p_ += 4;
        } else if (s_.compare(p_, 5, "false") == 0) {
            v.kind = JValue::Bool;
            p_ += 5;
        } else if (s_.compare(p_, 4, "null") == 0) {
            p_ += 4;
        } else {
            const char* begin = s_.c_str() + p_;
            char* end = nullptr;
            v.kind = JValue::Number;
            v.num = std::strtod(begin, &end);
            if (end == begin) err("bad value");
            p_ += static_cast<Size>(end - begin);
        }
        return v;
    }
    const std::string& s_;
    const char* what_;
    Size p_ = 0;
};

}  // namespace libaccint::detail
