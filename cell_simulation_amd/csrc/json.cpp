// json.cpp -- recursive-descent JSON reader (RFC 8259 subset sufficient for model files).
#include "json.hpp"

#include <cerrno>
#include <cstdlib>
#include <cstring>

#include "../../include/cell_sim.h"
#include "model.hpp"

namespace cs {
namespace json {

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string &what)
    {
        throw Error(CS_ERR_PARSE, "JSON: " + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r'))
            ++i;
    }
    bool lit(const char *w)
    {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string &out, uint32_t cp)
    {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4()
    {
        if (i + 4 > s.size())
            fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9')
                v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f')
                v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                v |= uint32_t(c - 'A' + 10);
            else
                fail("bad \\u escape");
        }
        return v;
    }
    std::string string()
    {
        if (s[i] != '"')
            fail("expected string");
        ++i;
        std::string out;
        while (true) {
            if (i >= s.size())
                fail("unterminated string");
            char c = s[i++];
            if (c == '"')
                break;
            if (static_cast<unsigned char>(c) < 0x20)
                fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size())
                fail("unterminated escape");
            char e = s[i++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                uint32_t cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                    i += 2;
                    uint32_t lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }
    Value number()
    {
        size_t start = i;
        bool integer = true, negative = false;
        if (s[i] == '-') {
            negative = true;
            ++i;
        }
        if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9'))
            fail("bad number");
        while (i < s.size() && s[i] >= '0' && s[i] <= '9')
            ++i;
        if (i < s.size() && s[i] == '.') {
            integer = false;
            ++i;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9'))
                fail("bad fraction");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9')
                ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            integer = false;
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-'))
                ++i;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9'))
                fail("bad exponent");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9')
                ++i;
        }
        std::string tok = s.substr(start, i - start);
        Value v;
        v.kind = Value::Number;
        v.num = std::strtod(tok.c_str(), nullptr);
        if (integer && !negative) {
            errno = 0;
            unsigned long long u = std::strtoull(tok.c_str(), nullptr, 10);
            if (errno == 0) {
                v.is_uint = true;
                v.uint_val = u;
            }
        }
        return v;
    }
    Value value(int depth)
    {
        if (depth > 256)
            fail("nesting too deep");
        ws();
        if (i >= s.size())
            fail("unexpected end");
        char c = s[i];
        Value v;
        if (c == '{') {
            ++i;
            v.kind = Value::Object;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            while (true) {
                ws();
                std::string key = string();
                ws();
                if (i >= s.size() || s[i] != ':')
                    fail("expected ':'");
                ++i;
                v.obj[key] = value(depth + 1);
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            ++i;
            v.kind = Value::Array;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
            v.b = false;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else {
            v = number();
        }
        return v;
    }
};

}  // namespace

Value parse(const std::string &text)
{
    Parser p{text};
    Value v = p.value(0);
    p.ws();
    if (p.i != text.size())
        p.fail("trailing characters");
    return v;
}

}  // namespace json
}  // namespace cs
