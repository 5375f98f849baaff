// json.hpp -- a small JSON reader for the rtm3d input files.
//
// Accepts standard JSON (objects, arrays, strings with escapes, numbers,
// true/false/null).  Numbers remember whether they were written as integers
// so that the parameter loader can apply the reference's type rules.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rtm3d {

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum class Type { Null, Bool, Int, Float, String, Array, Object };

    Json() = default;
    static Json parse(const std::string &text);
    static Json parse_file(const std::string &path);

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_number() const { return type_ == Type::Int || type_ == Type::Float; }
    bool has(const std::string &key) const;
    const Json &at(const std::string &key) const;  // throws JsonError if missing

    bool as_bool() const;
    double as_double() const;
    long long as_int() const;
    const std::string &as_string() const;
    const std::vector<Json> &as_array() const;
    const std::map<std::string, Json> &as_object() const;

    static const char *type_name(Type t);

private:
    Type type_ = Type::Null;
    bool b_ = false;
    double d_ = 0.0;
    long long i_ = 0;
    std::string s_;
    std::vector<Json> a_;
    std::map<std::string, Json> o_;

    friend class JsonParser;
};

}  // namespace rtm3d
