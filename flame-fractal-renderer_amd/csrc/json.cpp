// JSON parser and printer for flame files (see json.hpp).
#include "json.hpp"

#include <cerrno>
#include <clocale>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ffr {

class JsonParser {
public:
    explicit JsonParser(const std::string &t) : s_(t) {}

    Json document()
    {
        skip();
        Json v = value(0);
        skip();
        if (p_ != s_.size())
            fail("syntax error while parsing value - unexpected trailing input");
        return v;
    }

private:
    const std::string &s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &what)
    {
        size_t line = 1, col = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); ++i) {
            if (s_[i] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw JsonError("[json.exception.parse_error.101] parse error at line " +
                        std::to_string(line) + ", column " + std::to_string(col) +
                        ": " + what);
    }

    void skip()
    {
        for (;;) {
            while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' ||
                                      s_[p_] == '\n' || s_[p_] == '\r'))
                ++p_;
            if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '/') {
                while (p_ < s_.size() && s_[p_] != '\n')
                    ++p_;
                continue;
            }
            if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '*') {
                size_t e = s_.find("*/", p_ + 2);
                if (e == std::string::npos) {
                    p_ = s_.size();
                    fail("syntax error - unterminated comment");
                }
                p_ = e + 2;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == '/')
                fail("syntax error - invalid comment");
            return;
        }
    }

    Json value(int depth)
    {
        if (depth > 512)
            fail("syntax error - nesting too deep");
        if (p_ >= s_.size())
            fail("syntax error while parsing value - unexpected end of input");
        char c = s_[p_];
        Json v;
        if (c == '{') {
            ++p_;
            v.type_ = Json::Type::Object;
            skip();
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                skip();
                if (p_ >= s_.size() || s_[p_] != '"')
                    fail("syntax error while parsing object key - expected string");
                std::string key = string_lit();
                skip();
                if (p_ >= s_.size() || s_[p_] != ':')
                    fail("syntax error while parsing object separator - expected ':'");
                ++p_;
                skip();
                v.obj_[key] = value(depth + 1); // the last value of a key wins
                skip();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("syntax error while parsing object - expected '}'");
            }
        }
        if (c == '[') {
            ++p_;
            v.type_ = Json::Type::Array;
            skip();
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                skip();
                v.arr_.push_back(value(depth + 1));
                skip();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("syntax error while parsing array - expected ']'");
            }
        }
        if (c == '"') {
            v.type_ = Json::Type::String;
            v.s_ = string_lit();
            return v;
        }
        if (s_.compare(p_, 4, "true") == 0) {
            p_ += 4;
            v.type_ = Json::Type::Bool;
            v.b_ = true;
            return v;
        }
        if (s_.compare(p_, 5, "false") == 0) {
            p_ += 5;
            v.type_ = Json::Type::Bool;
            return v;
        }
        if (s_.compare(p_, 4, "null") == 0) {
            p_ += 4;
            return v;
        }
        if (c == '-' || (c >= '0' && c <= '9'))
            return number();
        fail("syntax error while parsing value - invalid literal");
    }

    static void put_utf8(std::string &o, uint32_t cp)
    {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }

    uint32_t hex4()
    {
        if (p_ + 4 > s_.size())
            fail("syntax error while parsing string - bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            char h = s_[p_++];
            v <<= 4;
            if (h >= '0' && h <= '9')
                v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f')
                v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F')
                v |= (uint32_t)(h - 'A' + 10);
            else
                fail("syntax error while parsing string - bad \\u escape");
        }
        return v;
    }

    std::string string_lit()
    {
        ++p_; // opening quote
        std::string o;
        for (;;) {
            if (p_ >= s_.size())
                fail("syntax error while parsing string - missing closing quote");
            unsigned char c = (unsigned char)s_[p_++];
            if (c == '"')
                return o;
            if (c < 0x20)
                fail("syntax error while parsing string - control character");
            if (c != '\\') {
                o += (char)c;
                continue;
            }
            if (p_ >= s_.size())
                fail("syntax error while parsing string - bad escape");
            char e = s_[p_++];
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
                if (cp >= 0xD800 && cp <= 0xDBFF) {
                    if (p_ + 1 < s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
                        p_ += 2;
                        uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF)
                            fail("syntax error while parsing string - bad surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else {
                        fail("syntax error while parsing string - bad surrogate");
                    }
                } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                    fail("syntax error while parsing string - bad surrogate");
                }
                put_utf8(o, cp);
                break;
            }
            default:
                fail("syntax error while parsing string - bad escape");
            }
        }
    }

    Json number()
    {
        size_t start = p_;
        bool neg = false, is_float = false;
        if (s_[p_] == '-') {
            neg = true;
            ++p_;
        }
        if (p_ >= s_.size() || !(s_[p_] >= '0' && s_[p_] <= '9'))
            fail("syntax error while parsing value - invalid number");
        if (s_[p_] == '0') {
            ++p_;
        } else {
            while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9')
                ++p_;
        }
        if (p_ < s_.size() && s_[p_] == '.') {
            is_float = true;
            ++p_;
            if (p_ >= s_.size() || !(s_[p_] >= '0' && s_[p_] <= '9'))
                fail("syntax error while parsing value - invalid number");
            while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9')
                ++p_;
        }
        if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
            is_float = true;
            ++p_;
            if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-'))
                ++p_;
            if (p_ >= s_.size() || !(s_[p_] >= '0' && s_[p_] <= '9'))
                fail("syntax error while parsing value - invalid number");
            while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9')
                ++p_;
        }
        std::string tok = s_.substr(start, p_ - start);
        Json v;
        if (!is_float) {
            errno = 0;
            if (neg) {
                long long x = std::strtoll(tok.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.type_ = Json::Type::Int;
                    v.i_ = x;
                    return v;
                }
            } else {
                unsigned long long x = std::strtoull(tok.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.type_ = Json::Type::UInt;
                    v.u_ = x;
                    return v;
                }
            }
        }
        v.type_ = Json::Type::Float;
        v.f_ = std::strtod(tok.c_str(), nullptr); // "C" locale digits only
        return v;
    }
};

Json Json::parse(const std::string &text)
{
    JsonParser p(text);
    return p.document();
}

size_t Json::size() const
{
    if (isArray())
        return arr_.size();
    if (isObject())
        return obj_.size();
    throw JsonError("Json::size(): type is not an array or object");
}

bool Json::boolValue() const
{
    if (!isBool())
        throw JsonError("Json::boolValue(): type is not bool");
    return b_;
}

int64_t Json::intValue() const
{
    if (type_ == Type::Int)
        return i_;
    if (type_ == Type::UInt)
        return (int64_t)u_;
    throw JsonError("Json::intValue(): type is not int");
}

double Json::floatValue() const
{
    if (type_ == Type::Int)
        return (double)i_;
    if (type_ == Type::UInt)
        return (double)u_;
    if (type_ == Type::Float)
        return f_;
    throw JsonError("Json::floatValue(): type is not float");
}

const std::string &Json::stringValue() const
{
    if (!isString())
        throw JsonError("Json::stringValue(): type is not string");
    return s_;
}

const std::vector<Json> &Json::arrayValue() const
{
    if (!isArray())
        throw JsonError("Json::arrayValue(): type is not array");
    return arr_;
}

const std::map<std::string, Json> &Json::objectValue() const
{
    if (!isObject())
        throw JsonError("Json::objectValue(): type is not object");
    return obj_;
}

bool Json::has(const std::string &key) const
{
    return isObject() && obj_.find(key) != obj_.end();
}

const Json &Json::operator[](const std::string &key) const
{
    if (!isObject())
        throw JsonError("Json::operator[](char*): not an object");
    auto it = obj_.find(key);
    if (it == obj_.end())
        throw JsonError("Json::operator[](char*): key does not exist: " + key);
    return it->second;
}

const Json &Json::operator[](size_t index) const
{
    if (!isArray())
        throw JsonError("Json::operator[](size_t): not an array");
    if (index >= arr_.size())
        throw JsonError("Json::operator[](size_t): array index out of range: " +
                        std::to_string(index));
    return arr_[index];
}

// shortest decimal digits that read back to v, laid out like nlohmann's
// serializer (digits[000].0 / dig.its / 0.[000]digits / d.igitse+XX)
std::string format_double(double v)
{
    if (!std::isfinite(v))
        return "null";
    std::string out;
    if (std::signbit(v)) {
        out += '-';
        v = -v;
    }
    if (v == 0.0)
        return out + "0.0";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v)
            break;
    }
    // buf = d[.ddd]e[+-]XX
    std::string digits;
    const char *q = buf;
    for (; *q && *q != 'e'; ++q)
        if (*q >= '0' && *q <= '9')
            digits += *q;
    int e10 = std::atoi(q + 1);
    while (digits.size() > 1 && digits.back() == '0')
        digits.pop_back();
    int k = (int)digits.size();
    int n = e10 + 1; // position of the decimal point
    const int min_exp = -4, max_exp = 15;
    if (k <= n && n <= max_exp) {
        out += digits + std::string((size_t)(n - k), '0') + ".0";
    } else if (0 < n && n <= max_exp) {
        out += digits.substr(0, (size_t)n) + "." + digits.substr((size_t)n);
    } else if (min_exp < n && n <= 0) {
        out += "0." + std::string((size_t)(-n), '0') + digits;
    } else {
        out += digits.substr(0, 1);
        if (k > 1)
            out += "." + digits.substr(1);
        int e = n - 1;
        out += 'e';
        out += e < 0 ? '-' : '+';
        int a = e < 0 ? -e : e;
        char eb[16];
        std::snprintf(eb, sizeof eb, "%02d", a);
        out += eb;
    }
    return out;
}

static void dump_string(std::string &o, const std::string &s)
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
                std::snprintf(b, sizeof b, "\\u%04x", c);
                o += b;
            } else {
                o += (char)c;
            }
        }
    }
    o += '"';
}

void Json::dump_to(std::string &o) const
{
    switch (type_) {
    case Type::Null: o += "null"; break;
    case Type::Bool: o += b_ ? "true" : "false"; break;
    case Type::Int: o += std::to_string(i_); break;
    case Type::UInt: o += std::to_string(u_); break;
    case Type::Float: o += format_double(f_); break;
    case Type::String: dump_string(o, s_); break;
    case Type::Array: {
        o += '[';
        bool first = true;
        for (const auto &e : arr_) {
            if (!first)
                o += ',';
            first = false;
            e.dump_to(o);
        }
        o += ']';
        break;
    }
    case Type::Object: {
        o += '{';
        bool first = true;
        for (const auto &kv : obj_) {
            if (!first)
                o += ',';
            first = false;
            dump_string(o, kv.first);
            o += ':';
            kv.second.dump_to(o);
        }
        o += '}';
        break;
    }
    }
}

std::string Json::dump() const
{
    std::string o;
    dump_to(o);
    return o;
}

} // namespace ffr
