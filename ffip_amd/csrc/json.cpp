#include "json.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ffip_amd {

class JsonParser {
public:
    explicit JsonParser(const std::string& s) : s_(s) {}

    Json parse_document()
    {
        Json v = value();
        ws();
        if (p_ != s_.size())
            fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what) const
    {
        throw std::runtime_error(std::string("JSON parse error at offset ") + std::to_string(p_) + ": " + what);
    }

    void ws()
    {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\t' || s_[p_] == '\r'))
            ++p_;
    }

    bool lit(const char* w)
    {
        size_t n = std::strlen(w);
        if (s_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Json value()
    {
        ws();
        if (p_ >= s_.size())
            fail("unexpected end");
        char c = s_[p_];
        Json v;
        if (c == '{') {
            v.kind_ = Json::Object;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (p_ >= s_.size() || s_[p_] != '"')
                    fail("expected key");
                std::string k = string();
                ws();
                if (p_ >= s_.size() || s_[p_] != ':')
                    fail("expected ':'");
                ++p_;
                Json item = value();
                v.obj_.emplace_back(std::move(k), std::move(item));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind_ = Json::Array;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.arr_.push_back(value());
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind_ = Json::String;
            v.str_ = string();
            return v;
        }
        if (lit("true")) {
            v.kind_ = Json::Bool;
            v.b_ = true;
            return v;
        }
        if (lit("false")) {
            v.kind_ = Json::Bool;
            return v;
        }
        if (lit("null"))
            return v;
        if (lit("NaN")) {
            v.kind_ = Json::Number;
            v.num_ = std::nan("");
            return v;
        }
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        double d = std::strtod(b, &e);
        if (e == b)
            fail("unexpected character");
        p_ += size_t(e - b);
        v.kind_ = Json::Number;
        v.num_ = d;
        return v;
    }

    std::string string()
    {
        ++p_;  // opening quote
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                if (p_ >= s_.size())
                    fail("bad escape");
                char e = s_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size())
                        fail("bad \\u escape");
                    unsigned cp = unsigned(std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16));
                    p_ += 4;
                    if (cp < 0x80)
                        out += char(cp);
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
        if (p_ >= s_.size())
            fail("unterminated string");
        ++p_;
        return out;
    }
};

Json Json::parse(const std::string& text)
{
    JsonParser p(text);
    return p.parse_document();
}

double Json::num() const
{
    if (kind_ == Bool)
        return b_ ? 1.0 : 0.0;
    if (kind_ != Number)
        throw std::runtime_error("JSON value is not a number");
    return num_;
}

int Json::integer() const
{
    return int(std::trunc(num()));
}

const std::string& Json::str() const
{
    if (kind_ != String)
        throw std::runtime_error("JSON value is not a string");
    return str_;
}

bool Json::boolean() const
{
    if (kind_ == Bool)
        return b_;
    return num() != 0;
}

size_t Json::size() const
{
    if (kind_ == Array)
        return arr_.size();
    if (kind_ == Object)
        return obj_.size();
    if (kind_ == Null)
        return 0;
    throw std::runtime_error("JSON value has no size");
}

const Json& Json::operator[](size_t i) const
{
    if (kind_ != Array || i >= arr_.size())
        throw std::runtime_error("JSON array index out of range");
    return arr_[i];
}

const Json* Json::find(const std::string& key) const
{
    if (kind_ != Object)
        return nullptr;
    for (auto& kv : obj_)
        if (kv.first == key)
            return &kv.second;
    return nullptr;
}

const Json& Json::at(const std::string& key) const
{
    const Json* j = find(key);
    if (!j)
        throw std::runtime_error("missing JSON key \"" + key + "\"");
    return *j;
}

double Json::num_or(const std::string& key, double dflt) const
{
    const Json* j = find(key);
    return j ? j->num() : dflt;
}

static void dump_string(const std::string& v, std::string& out)
{
    out += '"';
    for (unsigned char c : v) {
        if (c == '"' || c == '\\') {
            out += '\\';
            out += char(c);
        } else if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            out += buf;
        } else {
            out += char(c);
        }
    }
    out += '"';
}

static void dump_value(const Json& j, std::string& out)
{
    switch (j.kind()) {
    case Json::Null:
        out += "null";
        break;
    case Json::Bool:
        out += j.boolean() ? "true" : "false";
        break;
    case Json::Number: {
        double v = j.num();
        if (!std::isfinite(v))
            throw std::runtime_error("JSON cannot hold a non-finite number");
        char buf[32];
        std::snprintf(buf, sizeof buf, "%.17g", v);
        out += buf;
        break;
    }
    case Json::String:
        dump_string(j.str(), out);
        break;
    case Json::Array:
        out += '[';
        for (size_t i = 0; i < j.elems().size(); ++i) {
            if (i)
                out += ',';
            dump_value(j.elems()[i], out);
        }
        out += ']';
        break;
    case Json::Object:
        out += '{';
        for (size_t i = 0; i < j.items().size(); ++i) {
            if (i)
                out += ',';
            dump_string(j.items()[i].first, out);
            out += ':';
            dump_value(j.items()[i].second, out);
        }
        out += '}';
        break;
    }
}

std::string Json::dump() const
{
    std::string out;
    dump_value(*this, out);
    return out;
}

}  // namespace ffip_amd
