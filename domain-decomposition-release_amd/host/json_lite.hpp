// Small JSON reader for argfile.json: objects, arrays, strings, numbers,
// true/false/null, and C/C++ comments (the shipped argfiles use `//` lines).
// Accessors follow the conversions the argfile reader relies on
// (BASE/fdtd3d_base_1_interface.hpp:55-76, jsoncpp 1.9.4 Value::asInt /
// asFloat / asBool / asString): a missing key reads as null, null converts
// to 0 / 0.0 / false / "", numbers convert between int and float by C++
// conversion (truncation toward zero), and a string where a number or bool
// is expected is an error.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace ddxhost {

class JsonValue {
 public:
  enum class Kind { null, boolean, integer, real, string, array, object };
  Kind kind = Kind::null;
  bool b = false;
  long long i = 0;
  double r = 0.0;
  std::string s;
  std::vector<JsonValue> items;
  std::map<std::string, JsonValue> members;

  // member access; a missing member (or a non-object) yields null
  const JsonValue &operator[](const std::string &key) const;
  const JsonValue &operator[](size_t idx) const;
  size_t size() const { return kind == Kind::array ? items.size() : kind == Kind::object ? members.size() : 0; }
  bool is_null() const { return kind == Kind::null; }

  // conversions; on a type that cannot convert they throw std::runtime_error
  int asInt() const;
  float asFloat() const;
  double asDouble() const;
  bool asBool() const;
  std::string asString() const;
};

// Parses `text`; on a syntax error returns false and sets `err` (with the
// line number); `out` then holds what was parsed before the error.
bool json_parse(const std::string &text, JsonValue &out, std::string &err);

}  // namespace ddxhost
