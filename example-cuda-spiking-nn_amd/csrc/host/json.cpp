#include "json.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace bn {
namespace json {

const Value* Value::find(const std::string& key) const {
    if (kind != Object) return nullptr;
    // Last occurrence wins, as in common JSON libraries.
    const Value* hit = nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) hit = &kv.second;
    return hit;
}

double Value::as_double() const {
    switch (kind) {
        case Integer: return (double)i;
        case Unsigned: return (double)u;
        case Float: return f;
        default: throw std::runtime_error("JSON value is not a number");
    }
}

namespace {

struct Parser {
    const std::string& t;
    size_t p = 0;
    explicit Parser(const std::string& text) : t(text) {}

    [[noreturn]] void fail(const std::string& what) {
        throw std::runtime_error("JSON parse error: " + what + " at offset " + std::to_string(p));
    }
    void ws() {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
    }
    bool eat(char c) {
        ws();
        if (p < t.size() && t[p] == c) {
            ++p;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (t.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }

    Value value() {
        ws();
        if (p >= t.size()) fail("unexpected end of input");
        const char c = t[p];
        Value v;
        if (c == '{') {
            ++p;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                if (p >= t.size() || t[p] != '"') fail("expected object key");
                std::string k = str();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
            } while (eat(','));
            expect('}');
            return v;
        }
        if (c == '[') {
            ++p;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value());
            } while (eat(','));
            expect(']');
            return v;
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
            v.b = false;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) return num();
        fail(std::string("unexpected character '") + c + "'");
    }

    std::string str() {
        ++p;  // opening quote
        std::string out;
        while (true) {
            if (p >= t.size()) fail("unterminated string");
            char c = t[p++];
            if (c == '"') break;
            if (c == '\\') {
                if (p >= t.size()) fail("bad escape");
                char e = t[p++];
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
                        if (p + 4 > t.size()) fail("bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(t.substr(p, 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }

    Value num() {
        const size_t start = p;
        bool is_float = false;
        if (t[p] == '-') ++p;
        while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        if (p < t.size() && t[p] == '.') {
            is_float = true;
            ++p;
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            is_float = true;
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        const std::string text = t.substr(start, p - start);
        if (text == "-" || text.empty()) fail("bad number");
        Value v;
        if (!is_float) {
            errno = 0;
            if (text[0] == '-') {
                const long long x = std::strtoll(text.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.kind = Value::Integer;
                    v.i = x;
                    return v;
                }
            } else {
                const unsigned long long x = std::strtoull(text.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.kind = Value::Unsigned;
                    v.u = x;
                    v.i = (int64_t)x;
                    return v;
                }
            }
        }
        v.kind = Value::Float;
        v.f = std::strtod(text.c_str(), nullptr);
        return v;
    }
};

}  // namespace

Value parse(const std::string& text) {
    Parser ps(text);
    Value v = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.fail("trailing characters");
    return v;
}

std::string quote(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += c;
        } else if (c == '\n') {
            o += "\\n";
        } else {
            o += c;
        }
    }
    return o + "\"";
}

std::string number(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

}  // namespace json
}  // namespace bn
