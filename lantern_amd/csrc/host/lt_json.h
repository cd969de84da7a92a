/*
 * lt_json.h - a small JSON reader for Lantern scene files.
 *
 * The reference parses scenes with nlohmann::json and validates them against
 * scenes/scene.schema.json with json-schema-validator (scene.cpp:154-213).
 * This reader covers the JSON the scene format uses: objects, arrays, numbers,
 * strings (with \uXXXX escapes restricted to the BMP), true/false/null.
 */
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace lt_json {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;   /* keeps file order */

    bool is_object() const { return type == Object; }
    bool is_array() const { return type == Array; }
    bool is_number() const { return type == Number; }
    bool is_string() const { return type == String; }
    /* nlohmann's count(key) on an object: 0 or 1 */
    size_t count(const std::string &key) const;
    /* Member lookup; returns a Null value when absent. */
    const Value &operator[](const std::string &key) const;
    const Value &operator[](size_t i) const;
    size_t size() const { return type == Array ? arr.size() : type == Object ? obj.size() : 0; }
};

/* Parses text; on failure returns false and writes a message with the byte
 * offset into err. */
bool parse(const std::string &text, Value &out, std::string &err);
bool parse_file(const std::string &path, Value &out, std::string &err);

} // namespace lt_json
