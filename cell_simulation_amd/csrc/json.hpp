// json.hpp -- small JSON reader for the model format (parameters / regions / particles / reactions).
//
// Objects keep their keys sorted (std::map), which is the iteration order of serde_json's default Map
// (a BTreeMap; the reference does not enable `preserve_order`, Cargo.lock:1629-1637). The reaction row
// order depends on it (src/simulation.rs:321). A repeated key keeps its last value, as serde_json does.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace cs {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;          // every number, as f64 (serde_json as_f64)
    bool is_uint = false;      // integer literal >= 0 that fits u64 (serde_json as_u64)
    uint64_t uint_val = 0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    const Value *get(const std::string &key) const
    {
        if (kind != Object)
            return nullptr;
        auto it = obj.find(key);
        return it == obj.end() ? nullptr : &it->second;
    }
};

// Throws cs::Error(CS_ERR_PARSE) on malformed input.
Value parse(const std::string &text);

}  // namespace json
}  // namespace cs
