#include "json_lite.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace ddxhost {

static const JsonValue &null_value() {
  static const JsonValue v;
  return v;
}

const JsonValue &JsonValue::operator[](const std::string &key) const {
  if (kind != Kind::object) return null_value();
  auto it = members.find(key);
  return it == members.end() ? null_value() : it->second;
}

const JsonValue &JsonValue::operator[](size_t idx) const {
  if (kind != Kind::array || idx >= items.size()) return null_value();
  return items[idx];
}

int JsonValue::asInt() const {
  switch (kind) {
    case Kind::null: return 0;
    case Kind::boolean: return b ? 1 : 0;
    case Kind::integer:
      if (i < std::numeric_limits<int>::min() || i > std::numeric_limits<int>::max())
        throw std::runtime_error("JSON integer out of int range");
      return (int)i;
    case Kind::real:
      if (!(r >= (double)std::numeric_limits<int>::min() && r <= (double)std::numeric_limits<int>::max()))
        throw std::runtime_error("JSON real out of int range");
      return (int)r;
    default: throw std::runtime_error("JSON value is not convertible to int");
  }
}

double JsonValue::asDouble() const {
  switch (kind) {
    case Kind::null: return 0.0;
    case Kind::boolean: return b ? 1.0 : 0.0;
    case Kind::integer: return (double)i;
    case Kind::real: return r;
    default: throw std::runtime_error("JSON value is not convertible to double");
  }
}

float JsonValue::asFloat() const { return (float)asDouble(); }

bool JsonValue::asBool() const {
  switch (kind) {
    case Kind::null: return false;
    case Kind::boolean: return b;
    case Kind::integer: return i != 0;
    case Kind::real: return r != 0.0;
    default: throw std::runtime_error("JSON value is not convertible to bool");
  }
}

std::string JsonValue::asString() const {
  char buf[64];
  switch (kind) {
    case Kind::null: return "";
    case Kind::string: return s;
    case Kind::boolean: return b ? "true" : "false";
    case Kind::integer: snprintf(buf, sizeof buf, "%lld", i); return buf;
    case Kind::real: snprintf(buf, sizeof buf, "%.17g", r); return buf;
    default: throw std::runtime_error("JSON value is not convertible to string");
  }
}

namespace {

struct Parser {
  const std::string &t;
  size_t p = 0;
  std::string err;

  explicit Parser(const std::string &text) : t(text) {}

  int line() const {
    int n = 1;
    for (size_t k = 0; k < p && k < t.size(); k++) n += t[k] == '\n';
    return n;
  }
  bool fail(const std::string &m) {
    if (err.empty()) err = "line " + std::to_string(line()) + ": " + m;
    return false;
  }
  void skip() {
    for (;;) {
      while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) p++;
      if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '/') {
        while (p < t.size() && t[p] != '\n') p++;
      } else if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '*') {
        size_t e = t.find("*/", p + 2);
        p = e == std::string::npos ? t.size() : e + 2;
      } else {
        return;
      }
    }
  }
  bool literal(const char *w) {
    size_t n = strlen(w);
    if (t.compare(p, n, w) != 0) return false;
    p += n;
    return true;
  }
  bool string_(std::string &out) {
    if (p >= t.size() || t[p] != '"') return fail("expected string");
    p++;
    out.clear();
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p >= t.size()) return fail("unterminated escape");
      char e = t[p++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          if (p + 4 > t.size()) return fail("short \\u escape");
          unsigned cp = (unsigned)strtoul(t.substr(p, 4).c_str(), nullptr, 16);
          p += 4;
          // UTF-8 encode (basic multilingual plane)
          if (cp < 0x80) {
            out += (char)cp;
          } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
          } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          }
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p >= t.size()) return fail("unterminated string");
    p++;
    return true;
  }
  bool number(JsonValue &v) {
    size_t s = p;
    if (p < t.size() && (t[p] == '-' || t[p] == '+')) p++;
    bool real = false;
    while (p < t.size()) {
      char c = t[p];
      if (c >= '0' && c <= '9') {
        p++;
      } else if (c == '.' || c == 'e' || c == 'E' || ((c == '-' || c == '+') && (t[p - 1] == 'e' || t[p - 1] == 'E'))) {
        real = true;
        p++;
      } else {
        break;
      }
    }
    std::string tok = t.substr(s, p - s);
    if (tok.empty() || tok == "-" || tok == "+") return fail("expected value");
    char *end = nullptr;
    if (!real) {
      errno = 0;
      long long x = strtoll(tok.c_str(), &end, 10);
      if (errno == 0 && end && *end == 0) {
        v.kind = JsonValue::Kind::integer;
        v.i = x;
        return true;
      }
    }
    double d = strtod(tok.c_str(), &end);
    if (!end || *end != 0) return fail("bad number '" + tok + "'");
    v.kind = JsonValue::Kind::real;
    v.r = d;
    return true;
  }
  bool value(JsonValue &v, int depth) {
    if (depth > 64) return fail("nesting too deep");
    skip();
    if (p >= t.size()) return fail("unexpected end of input");
    char c = t[p];
    if (c == '{') {
      p++;
      v.kind = JsonValue::Kind::object;
      skip();
      if (p < t.size() && t[p] == '}') {
        p++;
        return true;
      }
      for (;;) {
        skip();
        std::string key;
        if (!string_(key)) return false;
        skip();
        if (p >= t.size() || t[p] != ':') return fail("expected ':'");
        p++;
        JsonValue item;
        bool ok = value(item, depth + 1);
        v.members[key] = item;
        if (!ok) return false;
        skip();
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        if (p < t.size() && t[p] == '}') {
          p++;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      p++;
      v.kind = JsonValue::Kind::array;
      skip();
      if (p < t.size() && t[p] == ']') {
        p++;
        return true;
      }
      for (;;) {
        JsonValue item;
        bool ok = value(item, depth + 1);
        v.items.push_back(item);
        if (!ok) return false;
        skip();
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        if (p < t.size() && t[p] == ']') {
          p++;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JsonValue::Kind::string;
      return string_(v.s);
    }
    if (literal("true")) {
      v.kind = JsonValue::Kind::boolean;
      v.b = true;
      return true;
    }
    if (literal("false")) {
      v.kind = JsonValue::Kind::boolean;
      v.b = false;
      return true;
    }
    if (literal("null")) {
      v.kind = JsonValue::Kind::null;
      return true;
    }
    return number(v);
  }
};

}  // namespace

bool json_parse(const std::string &text, JsonValue &out, std::string &err) {
  Parser ps(text);
  out = JsonValue();
  bool ok = ps.value(out, 0);
  if (ok) {
    ps.skip();
    if (ps.p != text.size()) ok = ps.fail("trailing characters after the root value");
  }
  err = ps.err;
  return ok;
}

}  // namespace ddxhost
