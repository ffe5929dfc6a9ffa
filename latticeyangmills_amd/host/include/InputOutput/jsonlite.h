// A small JSON reader for the run-description files of the reference's
// front-end (keys checked at src/InputOutput/JsonInput/*): objects, arrays,
// strings, numbers, true/false/null.  Parse errors throw std::runtime_error.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace LatticeIO {
struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool has(const std::string &k) const { return kind == Object && obj.count(k); }
    const JsonValue &operator[](const std::string &k) const;  // throws if missing
    double number() const;
    int integer() const;
    const std::string &string() const;

    static JsonValue parse(const std::string &text);
    static JsonValue parseFile(const std::string &path);
};
}  // namespace LatticeIO
