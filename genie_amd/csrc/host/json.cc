// Recursive-descent JSON reader (RFC 8259 grammar); see json.h.
#include "json.h"

#include <cerrno>
#include <cstdlib>
#include <cstring>

namespace genie_amd {
namespace {

constexpr int kMaxDepth = 256;

struct Parser {
    const std::string& t;
    size_t p = 0;
    std::string err;

    explicit Parser(const std::string& text) : t(text) {}

    bool fail(const std::string& what)
    {
        if (err.empty())
            err = what + " at offset " + std::to_string(p);
        return false;
    }
    void ws()
    {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r'))
            ++p;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (t.compare(p, n, w) != 0)
            return fail("unknown literal");
        p += n;
        return true;
    }
    static void utf8(std::string& s, uint32_t c)
    {
        if (c < 0x80) {
            s += (char)c;
        } else if (c < 0x800) {
            s += (char)(0xC0 | (c >> 6));
            s += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            s += (char)(0xE0 | (c >> 12));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        } else {
            s += (char)(0xF0 | (c >> 18));
            s += (char)(0x80 | ((c >> 12) & 0x3F));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        }
    }
    bool hex4(uint32_t& v)
    {
        if (p + 4 > t.size())
            return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = t[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool str(std::string& s)
    {
        if (p >= t.size() || t[p] != '"')
            return fail("expected a string");
        ++p;
        while (true) {
            if (p >= t.size())
                return fail("unterminated string");
            const char c = t[p++];
            if (c == '"')
                return true;
            if ((unsigned char)c < 0x20)
                return fail("control character in string");
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p >= t.size())
                return fail("unterminated escape");
            const char e = t[p++];
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
                uint32_t v;
                if (!hex4(v))
                    return false;
                if (v >= 0xD800 && v < 0xDC00) {  // surrogate pair
                    uint32_t w;
                    if (p + 2 > t.size() || t[p] != '\\' || t[p + 1] != 'u')
                        return fail("lone surrogate");
                    p += 2;
                    if (!hex4(w) || w < 0xDC00 || w >= 0xE000)
                        return fail("bad surrogate pair");
                    v = 0x10000 + ((v - 0xD800) << 10) + (w - 0xDC00);
                }
                utf8(s, v);
                break;
            }
            default: return fail("bad escape");
            }
        }
    }
    bool num(Json& j)
    {
        const size_t b = p;
        if (p < t.size() && t[p] == '-')
            ++p;
        if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9'))
            return fail("bad number");
        if (t[p] == '0')
            ++p;
        else
            while (p < t.size() && t[p] >= '0' && t[p] <= '9')
                ++p;
        bool real = false;
        if (p < t.size() && t[p] == '.') {
            real = true;
            ++p;
            if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9'))
                return fail("bad fraction");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9')
                ++p;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            real = true;
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-'))
                ++p;
            if (p >= t.size() || !(t[p] >= '0' && t[p] <= '9'))
                return fail("bad exponent");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9')
                ++p;
        }
        const std::string lit_s = t.substr(b, p - b);
        j.d = std::strtod(lit_s.c_str(), nullptr);
        if (real) {
            j.kind = Json::Double;
            return true;
        }
        errno = 0;
        const long long v = std::strtoll(lit_s.c_str(), nullptr, 10);
        j.i = v;
        j.kind = (errno == 0 && v >= INT32_MIN && v <= INT32_MAX) ? Json::Int : Json::BigInt;
        return true;
    }
    bool value(Json& j, int depth)
    {
        if (depth > kMaxDepth)
            return fail("nesting too deep");
        ws();
        if (p >= t.size())
            return fail("unexpected end");
        const char c = t[p];
        if (c == '{') {
            ++p;
            j.kind = Json::Object;
            ws();
            if (p < t.size() && t[p] == '}') {
                ++p;
                return true;
            }
            while (true) {
                ws();
                std::string key;
                if (!str(key))
                    return false;
                ws();
                if (p >= t.size() || t[p] != ':')
                    return fail("expected ':'");
                ++p;
                j.o.emplace_back(std::move(key), Json());
                if (!value(j.o.back().second, depth + 1))
                    return false;
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            j.kind = Json::Array;
            ws();
            if (p < t.size() && t[p] == ']') {
                ++p;
                return true;
            }
            while (true) {
                j.a.emplace_back();
                if (!value(j.a.back(), depth + 1))
                    return false;
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            j.kind = Json::String;
            return str(j.s);
        }
        if (c == 't') {
            j.kind = Json::Bool;
            j.b = true;
            return lit("true");
        }
        if (c == 'f') {
            j.kind = Json::Bool;
            j.b = false;
            return lit("false");
        }
        if (c == 'n') {
            j.kind = Json::Null;
            return lit("null");
        }
        return num(j);
    }
};

}  // namespace

bool parse_json(const std::string& text, Json& out, std::string* error)
{
    Parser ps(text);
    out = Json();
    if (!ps.value(out, 0)) {
        if (error)
            *error = ps.err;
        return false;
    }
    ps.ws();
    if (ps.p != text.size()) {
        ps.fail("trailing characters");
        if (error)
            *error = ps.err;
        return false;
    }
    return true;
}

}  // namespace genie_amd
