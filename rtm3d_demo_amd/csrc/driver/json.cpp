// json.cpp -- recursive-descent JSON reader (see json.hpp).
#include "json.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace rtm3d {

class JsonParser {
public:
    explicit JsonParser(const std::string &t) : t_(t) {}

    Json document()
    {
        Json v = value();
        ws();
        if (p_ != t_.size()) error("trailing characters");
        return v;
    }

private:
    const std::string &t_;
    size_t p_ = 0;

    [[noreturn]] void error(const char *what)
    {
        size_t line = 1;
        for (size_t i = 0; i < p_ && i < t_.size(); ++i) line += t_[i] == '\n';
        std::ostringstream os;
        os << "JSON parse error at line " << line << ": " << what;
        throw JsonError(os.str());
    }

    void ws()
    {
        while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) ++p_;
    }

    bool lit(const char *s)
    {
        size_t n = 0;
        while (s[n]) ++n;
        if (t_.compare(p_, n, s) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Json value()
    {
        ws();
        if (p_ >= t_.size()) error("unexpected end of input");
        Json v;
        const char c = t_[p_];
        if (c == '{') {
            v.type_ = Json::Type::Object;
            ++p_;
            ws();
            if (p_ < t_.size() && t_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (p_ >= t_.size() || t_[p_] != '"') error("expected a key string");
                std::string k = string();
                ws();
                if (p_ >= t_.size() || t_[p_] != ':') error("expected ':'");
                ++p_;
                v.o_[k] = value();
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == '}') {
                    ++p_;
                    return v;
                }
                error("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type_ = Json::Type::Array;
            ++p_;
            ws();
            if (p_ < t_.size() && t_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.a_.push_back(value());
                ws();
                if (p_ < t_.size() && t_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < t_.size() && t_[p_] == ']') {
                    ++p_;
                    return v;
                }
                error("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type_ = Json::Type::String;
            v.s_ = string();
            return v;
        }
        if (lit("true")) {
            v.type_ = Json::Type::Bool;
            v.b_ = true;
            return v;
        }
        if (lit("false")) {
            v.type_ = Json::Type::Bool;
            v.b_ = false;
            return v;
        }
        if (lit("null")) return v;
        return number();
    }

    std::string string()
    {
        ++p_;  // opening quote
        std::string s;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= t_.size()) error("bad escape");
            char e = t_[p_++];
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
                if (p_ + 4 > t_.size()) error("bad \\u escape");
                unsigned cp = (unsigned)std::strtoul(t_.substr(p_, 4).c_str(), nullptr, 16);
                p_ += 4;
                if (cp < 0x80) {
                    s += (char)cp;
                } else if (cp < 0x800) {
                    s += (char)(0xC0 | (cp >> 6));
                    s += (char)(0x80 | (cp & 0x3F));
                } else {
                    s += (char)(0xE0 | (cp >> 12));
                    s += (char)(0x80 | ((cp >> 6) & 0x3F));
                    s += (char)(0x80 | (cp & 0x3F));
                }
                break;
            }
            default: error("bad escape");
            }
        }
        if (p_ >= t_.size()) error("unterminated string");
        ++p_;
        return s;
    }

    Json number()
    {
        const size_t start = p_;
        bool is_float = false;
        if (t_[p_] == '-' || t_[p_] == '+') ++p_;
        while (p_ < t_.size()) {
            const char c = t_[p_];
            if (std::isdigit((unsigned char)c)) {
                ++p_;
            } else if (c == '.' || c == 'e' || c == 'E' ||
                       ((c == '-' || c == '+') && (t_[p_ - 1] == 'e' || t_[p_ - 1] == 'E'))) {
                is_float = true;
                ++p_;
            } else {
                break;
            }
        }
        if (p_ == start) error("unexpected character");
        const std::string tok = t_.substr(start, p_ - start);
        char *end = nullptr;
        Json v;
        if (is_float) {
            v.type_ = Json::Type::Float;
            v.d_ = std::strtod(tok.c_str(), &end);
            v.i_ = (long long)v.d_;
        } else {
            v.type_ = Json::Type::Int;
            v.i_ = std::strtoll(tok.c_str(), &end, 10);
            v.d_ = (double)v.i_;
        }
        if (!end || *end) error("malformed number");
        return v;
    }
};

Json Json::parse(const std::string &text)
{
    JsonParser p(text);
    return p.document();
}

Json Json::parse_file(const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw JsonError("[ Input File '" + path + "' doesn't exists. Please check input file path. ]");
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

bool Json::has(const std::string &key) const
{
    return type_ == Type::Object && o_.count(key) != 0;
}

const Json &Json::at(const std::string &key) const
{
    if (type_ != Type::Object) throw JsonError("not a JSON object");
    auto it = o_.find(key);
    if (it == o_.end()) throw JsonError("missing key '" + key + "'");
    return it->second;
}

const char *Json::type_name(Type t)
{
    switch (t) {
    case Type::Null: return "Null";
    case Type::Bool: return "Boolean";
    case Type::Int: return "Integer";
    case Type::Float: return "Float";
    case Type::String: return "String";
    case Type::Array: return "Array";
    case Type::Object: return "Object";
    }
    return "?";
}

bool Json::as_bool() const
{
    if (type_ != Type::Bool) throw JsonError(std::string("type must be boolean, but is ") + type_name(type_));
    return b_;
}

double Json::as_double() const
{
    if (!is_number()) throw JsonError(std::string("type must be number, but is ") + type_name(type_));
    return d_;
}

long long Json::as_int() const
{
    if (!is_number()) throw JsonError(std::string("type must be number, but is ") + type_name(type_));
    return type_ == Type::Int ? i_ : (long long)d_;
}

const std::string &Json::as_string() const
{
    if (type_ != Type::String) throw JsonError(std::string("type must be string, but is ") + type_name(type_));
    return s_;
}

const std::vector<Json> &Json::as_array() const
{
    if (type_ != Type::Array) throw JsonError("type must be array");
    return a_;
}

const std::map<std::string, Json> &Json::as_object() const
{
    if (type_ != Type::Object) throw JsonError("type must be object");
    return o_;
}

}  // namespace rtm3d
