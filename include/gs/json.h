// A JSON document as a tree of values (the role of engine/include/gs/json_parse.h
// for transforms.json): objects keep their keys in file order.
#pragma once

#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace gs {

// One value; `type` names the member that holds it (the names of the
// reference's DOM, so code written against it reads these unchanged).
struct JsonValue {
  enum class Type { Null, Bool, Number, String, Array, Object };

  Type type = Type::Null;
  std::vector<std::pair<std::string, JsonValue>> object;  // members in file order
  std::vector<JsonValue> array;
  std::string string;
  double number = 0.0;
  bool boolean = false;

  // The member `key` of an object (the last one if repeated), else nullptr.
  const JsonValue* find(const std::string& key) const;
  double number_or(const std::string& key, double fallback) const;
  std::string string_or(const std::string& key, const std::string& fallback) const;
};

// Parses RFC 8259 JSON; throws std::runtime_error naming the byte offset of
// the first error.
JsonValue parse_json(std::string_view text);

}  // namespace gs
