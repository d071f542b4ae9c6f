// Minimal JSON reader for the configuration documents ffip's Python front end
// writes (ffip/simulation.py:362-402).  Numbers are parsed with strtod, so the
// doubles are the ones the writer serialised.  Objects keep insertion order.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ffip_amd {

class Json {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json parse(const std::string& text);

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Null; }
    bool is_number() const { return kind_ == Number; }
    bool is_string() const { return kind_ == String; }
    bool is_array() const { return kind_ == Array; }
    bool is_object() const { return kind_ == Object; }

    double num() const;
    int integer() const;  // truncating conversion, as nlohmann get<int>()
    const std::string& str() const;
    bool boolean() const;

    size_t size() const;
    const Json& operator[](size_t i) const;
    const Json& at(const std::string& key) const;  // throws "missing key"
    const Json* find(const std::string& key) const;
    bool has(const std::string& key) const { return find(key) != nullptr; }
    const std::vector<std::pair<std::string, Json>>& items() const { return obj_; }
    const std::vector<Json>& elems() const { return arr_; }

    double num_or(const std::string& key, double dflt) const;

    // Serialises the value (numbers with 17 significant digits, so they parse
    // back to the same doubles).
    std::string dump() const;

private:
    Kind kind_ = Null;
    double num_ = 0;
    bool b_ = false;
    std::string str_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;

    friend class JsonParser;
};

}  // namespace ffip_amd
