// json.hpp -- a small JSON DOM: parser, and a printer in nlohmann's dump(4)
// layout (object keys in byte order, 4-space indent, ": " separators,
// integers kept as integers, doubles as their shortest round-trip decimal
// with ".0" when integral).  Used by the baked-light glTF writer/reader
// (format.cpp) and the glTF scene loader (gltf.cpp).
#pragma once

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace vlb {
namespace json {

struct JVal {
    enum Kind { Null, Bool, Int, UInt, Float, Str, Arr, Obj } kind = Null;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;
    double f = 0.0;
    std::string s;
    std::vector<JVal> a;
    std::map<std::string, JVal> o;
};

struct Parser {
    const char* p;
    const char* end;
    std::string err;
    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char* w)
    {
        size_t n = strlen(w);
        if ((size_t)(end - p) < n || strncmp(p, w, n) != 0) return false;
        p += n;
        return true;
    }
    static void put_utf8(std::string& s, uint32_t cp)
    {
        if (cp < 0x80) s += (char)cp;
        else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(uint32_t& v)
    {
        if (end - p < 4) return false;
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool str(std::string& s)
    {
        if (p >= end || *p != '"') return fail("expected string");
        ++p;
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= end) return fail("bad escape");
                char e = *p++;
                switch (e) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return fail("bad \\u escape");
                    if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t lo;
                        if (!hex4(lo)) return fail("bad surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(s, cp);
                    break;
                }
                default: return fail("bad escape");
                }
            } else {
                s += c;
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool num(JVal& v)
    {
        const char* s = p;
        bool neg = false, isfloat = false;
        if (p < end && *p == '-') {
            neg = true;
            ++p;
        }
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                           *p == '+' || *p == '-')) {
            if (*p == '.' || *p == 'e' || *p == 'E') isfloat = true;
            ++p;
        }
        std::string t(s, p);
        if (t.empty() || t == "-") return fail("bad number");
        if (!isfloat) {
            errno = 0;
            if (neg) {
                long long x = strtoll(t.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.kind = JVal::Int;
                    v.i = x;
                    return true;
                }
            } else {
                unsigned long long x = strtoull(t.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.kind = JVal::UInt;
                    v.u = x;
                    return true;
                }
            }
        }
        v.kind = JVal::Float;
        v.f = strtod(t.c_str(), nullptr);
        return true;
    }
    bool value(JVal& v, int depth)
    {
        if (depth > 512) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        char c = *p;
        if (c == '{') {
            ++p;
            v.kind = JVal::Obj;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.o[k] = std::move(x);
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            v.kind = JVal::Arr;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return true;
            }
            for (;;) {
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.a.push_back(std::move(x));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return str(v.s);
        }
        if (lit("true")) {
            v.kind = JVal::Bool;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JVal::Bool;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.kind = JVal::Null;
            return true;
        }
        return num(v);
    }
};

// Shortest decimal that round-trips through strtod, in nlohmann's layout:
// plain notation for decimal exponents in [-4, 14], otherwise d.ddde+XX.
inline std::string fmt_double(double x)
{
    if (!std::isfinite(x)) return "null";
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        snprintf(buf, sizeof buf, "%.*e", prec - 1, x);
        if (strtod(buf, nullptr) == x) break;
    }
    // split mantissa digits and exponent
    std::string m(buf);
    size_t epos = m.find('e');
    int e10 = atoi(m.c_str() + epos + 1);
    std::string digits;
    bool neg = m[0] == '-';
    for (size_t i = neg ? 1 : 0; i < epos; ++i)
        if (m[i] != '.') digits += m[i];
    std::string out = neg ? "-" : "";
    const int n = (int)digits.size();
    if (e10 >= -4 && e10 <= 14) {
        if (e10 >= 0) {
            if (n <= e10 + 1) {
                out += digits + std::string(e10 + 1 - n, '0') + ".0";
            } else {
                out += digits.substr(0, e10 + 1) + "." + digits.substr(e10 + 1);
            }
        } else {
            out += "0." + std::string(-e10 - 1, '0') + digits;
        }
    } else {
        out += digits.substr(0, 1);
        if (n > 1) out += "." + digits.substr(1);
        char eb[16];
        snprintf(eb, sizeof eb, "e%c%02d", e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
        out += eb;
    }
    return out;
}

inline void esc(std::string& o, const std::string& s)
{
    o += '"';
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\b': o += "\\b"; break;
        case '\f': o += "\\f"; break;
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
    o += '"';
}

inline void dump(std::string& o, const JVal& v, int ind)
{
    switch (v.kind) {
    case JVal::Null: o += "null"; break;
    case JVal::Bool: o += v.b ? "true" : "false"; break;
    case JVal::Int: o += std::to_string(v.i); break;
    case JVal::UInt: o += std::to_string(v.u); break;
    case JVal::Float: o += fmt_double(v.f); break;
    case JVal::Str: esc(o, v.s); break;
    case JVal::Arr:
        if (v.a.empty()) {
            o += "[]";
            break;
        }
        o += "[\n";
        for (size_t i = 0; i < v.a.size(); ++i) {
            o += std::string(ind + 4, ' ');
            dump(o, v.a[i], ind + 4);
            o += i + 1 < v.a.size() ? ",\n" : "\n";
        }
        o += std::string(ind, ' ') + "]";
        break;
    case JVal::Obj: {
        if (v.o.empty()) {
            o += "{}";
            break;
        }
        o += "{\n";
        size_t i = 0;
        for (const auto& kv : v.o) {
            o += std::string(ind + 4, ' ');
            esc(o, kv.first);
            o += ": ";
            dump(o, kv.second, ind + 4);
            o += ++i < v.o.size() ? ",\n" : "\n";
        }
        o += std::string(ind, ' ') + "}";
        break;
    }
    }
}

inline JVal juint(uint64_t x)
{
    JVal v;
    v.kind = JVal::UInt;
    v.u = x;
    return v;
}
inline JVal jfloat(double x)
{
    JVal v;
    v.kind = JVal::Float;
    v.f = x;
    return v;
}
inline JVal jstr(std::string s)
{
    JVal v;
    v.kind = JVal::Str;
    v.s = std::move(s);
    return v;
}


// Parses a whole document; false on a syntax error or trailing text.
inline bool parse(const char* text, size_t len, JVal& root)
{
    Parser ps{text, text + len, {}};
    if (!ps.value(root, 0)) return false;
    ps.ws();
    return ps.p == ps.end;
}

}  // namespace json
}  // namespace vlb
