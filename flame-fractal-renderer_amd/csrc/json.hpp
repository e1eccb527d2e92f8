// Read-only JSON document model for flame files.
//
// Replaces the reference's wrapper over nlohmann::json (src/utils/json.hpp:44-105,
// src/utils/json.cpp): comments are accepted (the reference parses with
// ignore_comments = true, json.cpp:15-23), integers that fit are kept as
// integers and everything else as doubles (correctly rounded, as strtod),
// objects keep one value per key (the last one written) in key order, and the
// typed accessors throw JsonError with the reference's texts.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ffr {

class JsonError : public std::runtime_error {
public:
    explicit JsonError(const std::string &m) : std::runtime_error(m) {}
};

class Json {
public:
    enum class Type { Null, Bool, Int, UInt, Float, String, Array, Object };

    Json() = default;
    static Json parse(const std::string &text); // throws JsonError

    Type type() const { return type_; }
    bool isNull() const { return type_ == Type::Null; }
    bool isBool() const { return type_ == Type::Bool; }
    bool isInt() const { return type_ == Type::Int || type_ == Type::UInt; }
    bool isFloat() const { return isInt() || type_ == Type::Float; }
    bool isString() const { return type_ == Type::String; }
    bool isArray() const { return type_ == Type::Array; }
    bool isObject() const { return type_ == Type::Object; }

    size_t size() const;
    bool boolValue() const;
    int64_t intValue() const;
    double floatValue() const;
    const std::string &stringValue() const;
    const std::vector<Json> &arrayValue() const;
    const std::map<std::string, Json> &objectValue() const;
    bool has(const std::string &key) const;
    const Json &operator[](const std::string &key) const;
    const Json &operator[](size_t index) const;

    // compact text in nlohmann::json::dump() form (keys sorted, shortest
    // round-trip doubles with ".0" on integral values)
    std::string dump() const;

private:
    friend class JsonParser;
    Type type_ = Type::Null;
    bool b_ = false;
    int64_t i_ = 0;
    uint64_t u_ = 0;
    double f_ = 0.0;
    std::string s_;
    std::vector<Json> arr_;
    std::map<std::string, Json> obj_;
    void dump_to(std::string &out) const;
};

std::string format_double(double v); // nlohmann-style number text

} // namespace ffr
