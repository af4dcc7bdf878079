// Minimal JSON reader for the simulator's flat config files (objects, arrays, strings,
// numbers, booleans, null).  Numbers keep their integer/float nature the way a JSON library
// reports it, because the config loader converts them to float/uint32/int like the
// reference's `j.value(key, default)` (cpp_src/core/config.cpp:84-289).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace bn {
namespace json {

struct Value {
    enum Kind { Null, Bool, Integer, Unsigned, Float, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;
    double f = 0.0;
    std::string s;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // insertion order

    bool is_number() const { return kind == Integer || kind == Unsigned || kind == Float; }
    bool is_integer() const { return kind == Integer || kind == Unsigned; }
    const Value* find(const std::string& key) const;
    bool contains(const std::string& key) const { return find(key) != nullptr; }
    double as_double() const;
};

// Throws std::runtime_error("JSON parse error: ...") on malformed text.
Value parse(const std::string& text);

// Serialisation helpers for dumps.
std::string quote(const std::string& s);
std::string number(double v);

}  // namespace json
}  // namespace bn
