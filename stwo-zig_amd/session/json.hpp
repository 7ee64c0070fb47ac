// Strict JSON (RFC 8259) reader and a small ordered writer for the prover session's JSONL frames
// and reports.  The reader rejects what a strict frame parser must: trailing text, duplicate
// keys, control characters in strings, leading zeros, lone surrogates, nesting beyond 64 levels.
// Integers without a fraction or exponent keep their exact value (up to 2^64 - 1, or -2^63).
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sj {

struct ParseError : std::runtime_error {
    explicit ParseError(const std::string& m) : std::runtime_error(m) {}
};

struct Value {
    enum Kind { Null, Bool, Int, Float, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    bool neg = false;  // Int: value = neg ? -mag : mag
    uint64_t mag = 0;
    double d = 0;
    std::string s;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // in document order

    const Value* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    bool is_uint() const { return kind == Int && !neg; }
};

namespace detail {

struct Reader {
    const char* p;
    const char* end;
    int depth = 0;

    [[noreturn]] void fail(const char* what) { throw ParseError(what); }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool lit(const char* w) {
        size_t n = strlen(w);
        if ((size_t)(end - p) < n || memcmp(p, w, n) != 0) return false;
        p += n;
        return true;
    }
    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 63));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 63));
            o += (char)(0x80 | (cp & 63));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 63));
            o += (char)(0x80 | ((cp >> 6) & 63));
            o += (char)(0x80 | (cp & 63));
        }
    }
    uint32_t hex4() {
        if (end - p < 4) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        if (p >= end || *p != '"') fail("expected a string");
        p++;
        std::string o;
        for (;;) {
            if (p >= end) fail("unterminated string");
            unsigned char c = (unsigned char)*p++;
            if (c == '"') return o;
            if (c < 0x20) fail("control character in a string");
            if (c != '\\') {
                o += (char)c;
                continue;
            }
            if (p >= end) fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        if (!lit("\\u")) fail("lone surrogate");
                        uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("lone surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("lone surrogate");
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    Value number() {
        const char* s0 = p;
        bool neg = false;
        if (*p == '-') {
            neg = true;
            p++;
        }
        if (p >= end || *p < '0' || *p > '9') fail("bad number");
        if (*p == '0' && p + 1 < end && p[1] >= '0' && p[1] <= '9') fail("leading zero");
        bool integral = true;
        uint64_t mag = 0;
        bool overflow = false;
        while (p < end && *p >= '0' && *p <= '9') {
            uint64_t d = (uint64_t)(*p - '0');
            if (mag > (UINT64_MAX - d) / 10) overflow = true;
            mag = mag * 10 + d;
            p++;
        }
        if (p < end && *p == '.') {
            integral = false;
            p++;
            if (p >= end || *p < '0' || *p > '9') fail("bad fraction");
            while (p < end && *p >= '0' && *p <= '9') p++;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            integral = false;
            p++;
            if (p < end && (*p == '+' || *p == '-')) p++;
            if (p >= end || *p < '0' || *p > '9') fail("bad exponent");
            while (p < end && *p >= '0' && *p <= '9') p++;
        }
        Value v;
        if (integral) {
            if (overflow || (neg && mag > (1ull << 63))) fail("integer out of range");
            v.kind = Value::Int;
            v.neg = neg && mag != 0;
            v.mag = mag;
            v.d = neg ? -(double)mag : (double)mag;
        } else {
            v.kind = Value::Float;
            v.d = strtod(std::string(s0, p).c_str(), nullptr);
            if (!std::isfinite(v.d)) fail("number out of range");
        }
        return v;
    }
    Value value() {
        ws();
        if (p >= end) fail("unexpected end of input");
        Value v;
        char c = *p;
        if (c == '{' || c == '[') {
            if (++depth > 64) fail("nesting too deep");
            p++;
            ws();
            char close = c == '{' ? '}' : ']';
            v.kind = c == '{' ? Value::Object : Value::Array;
            if (p < end && *p == close) {
                p++;
                depth--;
                return v;
            }
            for (;;) {
                if (v.kind == Value::Object) {
                    ws();
                    std::string k = str();
                    if (v.get(k)) fail("duplicate key");
                    ws();
                    if (p >= end || *p != ':') fail("expected ':'");
                    p++;
                    Value item = value();
                    v.obj.emplace_back(std::move(k), std::move(item));
                } else {
                    v.arr.push_back(value());
                }
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == close) {
                    p++;
                    depth--;
                    return v;
                }
                fail("expected ',' or a closing bracket");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.s = str();
            return v;
        }
        if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("unexpected character");
    }
};

}  // namespace detail

inline Value parse(const std::string& text) {
    detail::Reader r{text.data(), text.data() + text.size()};
    Value v = r.value();
    r.ws();
    if (r.p != r.end) throw ParseError("trailing text after the JSON value");
    return v;
}

inline std::string quote(const std::string& s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    snprintf(b, sizeof b, "\\u%04x", c);
                    o += b;
                } else {
                    o += (char)c;
                }
        }
    }
    return o + "\"";
}

inline std::string num(double x) {
    if (!std::isfinite(x)) return "null";
    char b[40];
    snprintf(b, sizeof b, "%.17g", x);
    // keep a float a float for readers that distinguish (1 -> 1.0)
    if (!strpbrk(b, ".eEn")) strcat(b, ".0");
    return b;
}
inline std::string num(uint64_t x) { return std::to_string((unsigned long long)x); }

// Ordered object / array builders; values are already-encoded JSON text.
struct Obj {
    std::string s = "{";
    Obj& raw(const char* k, const std::string& v) {
        if (s.size() > 1) s += ',';
        s += quote(k);
        s += ':';
        s += v;
        return *this;
    }
    Obj& str(const char* k, const std::string& v) { return raw(k, quote(v)); }
    Obj& u(const char* k, uint64_t v) { return raw(k, num(v)); }
    Obj& f(const char* k, double v) { return raw(k, num(v)); }
    Obj& b(const char* k, bool v) { return raw(k, v ? "true" : "false"); }
    Obj& null(const char* k) { return raw(k, "null"); }
    std::string done() const { return s + "}"; }
};
struct Arr {
    std::string s = "[";
    Arr& raw(const std::string& v) {
        if (s.size() > 1) s += ',';
        s += v;
        return *this;
    }
    std::string done() const { return s + "]"; }
};

}  // namespace sj
