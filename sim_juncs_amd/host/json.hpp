// json.hpp -- minimal JSON reader for scene files (objects, arrays, numbers, strings, true,
// false, null).  Numbers beyond the double range (e.g. 1e999) read as +-infinity.
#pragma once

#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sjfh {
namespace json {

struct J {
    enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
    bool b = false;
    double x = 0;
    std::string s;
    std::vector<J> a;
    std::map<std::string, J> o;

    bool has(const std::string& k) const { return type == Obj && o.count(k); }
    const J& at(const std::string& k) const {
        auto it = o.find(k);
        if (type != Obj || it == o.end()) throw std::runtime_error("json: missing key '" + k + "'");
        return it->second;
    }
    const J& operator[](size_t i) const {
        if (type != Arr || i >= a.size()) throw std::runtime_error("json: index out of range");
        return a[i];
    }
    double num() const {
        if (type != Num) throw std::runtime_error("json: number expected");
        return x;
    }
};

class Parser {
public:
    explicit Parser(const std::string& t) : t_(t) {}
    J parse() {
        J v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* m) { throw std::runtime_error(std::string("json: ") + m + " at " + std::to_string(p_)); }
    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\t' || t_[p_] == '\r')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (t_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    J value() {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        J v;
        const char c = t_[p_];
        if (c == '{') {
            v.type = J::Obj;
            ++p_;
            ws();
            if (t_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (t_[p_] != '"') fail("key expected");
                std::string k = str();
                ws();
                if (t_[p_++] != ':') fail("':' expected");
                v.o[k] = value();
                ws();
                if (t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (t_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            v.type = J::Arr;
            ++p_;
            ws();
            if (t_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.a.push_back(value());
                ws();
                if (t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (t_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("',' or ']' expected");
            }
        }
        if (c == '"') {
            v.type = J::Str;
            v.s = str();
            return v;
        }
        if (lit("true")) {
            v.type = J::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.type = J::Bool;
            return v;
        }
        if (lit("null")) return v;
        const char* s = t_.c_str() + p_;
        char* end = nullptr;
        v.x = std::strtod(s, &end);
        if (end == s) fail("value expected");
        p_ += (size_t)(end - s);
        v.type = J::Num;
        return v;
    }
    std::string str() {
        std::string o;
        ++p_;
        while (p_ < t_.size() && t_[p_] != '"') {
            if (t_[p_] == '\\' && p_ + 1 < t_.size()) {
                ++p_;
                const char e = t_[p_];
                o += e == 'n' ? '\n' : e == 't' ? '\t' : e;
            } else {
                o += t_[p_];
            }
            ++p_;
        }
        if (p_ >= t_.size()) fail("unterminated string");
        ++p_;
        return o;
    }
};

inline J parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace sjfh
