// Recursive-descent JSON reader (RFC 8259): numbers through strtod on the
// token, strings with every escape (surrogate pairs to UTF-8).
#include "gs/json.h"

#include <cstdint>
#include <cstdlib>
#include <stdexcept>

namespace gs {
namespace {

class Parser {
 public:
  explicit Parser(std::string_view t) : t_(t) {}

  JsonValue document() {
    JsonValue v = value(0);
    space();
    if (i_ != t_.size()) error("trailing characters");
    return v;
  }

 private:
  std::string_view t_;
  size_t i_ = 0;

  [[noreturn]] void error(const char* what) const {
    throw std::runtime_error("JSON: " + std::string(what) + " at byte " + std::to_string(i_));
  }
  void space() {
    while (i_ < t_.size() && (t_[i_] == ' ' || t_[i_] == '\t' || t_[i_] == '\n' || t_[i_] == '\r')) ++i_;
  }
  bool eat(char c) {
    space();
    if (i_ < t_.size() && t_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) error((std::string("expected '") + c + "'").c_str());
  }
  bool word(std::string_view w) {
    if (t_.substr(i_, w.size()) == w) {
      i_ += w.size();
      return true;
    }
    return false;
  }

  JsonValue value(int depth) {
    if (depth > 256) error("nesting too deep");
    space();
    if (i_ >= t_.size()) error("unexpected end");
    JsonValue v;
    const char c = t_[i_];
    if (c == '{') {
      ++i_;
      v.type = JsonValue::Type::Object;
      if (eat('}')) return v;
      do {
        space();
        if (i_ >= t_.size() || t_[i_] != '"') error("expected a key");
        std::string key = string();
        expect(':');
        v.object.emplace_back(std::move(key), value(depth + 1));
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++i_;
      v.type = JsonValue::Type::Array;
      if (eat(']')) return v;
      do {
        v.array.push_back(value(depth + 1));
      } while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.type = JsonValue::Type::String;
      v.string = string();
    } else if (word("true")) {
      v.type = JsonValue::Type::Bool;
      v.boolean = true;
    } else if (word("false")) {
      v.type = JsonValue::Type::Bool;
    } else if (word("null")) {
      v.type = JsonValue::Type::Null;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      v.type = JsonValue::Type::Number;
      v.number = number();
    } else {
      error("unexpected character");
    }
    return v;
  }

  double number() {
    const size_t start = i_;
    if (t_[i_] == '-') ++i_;
    auto digits = [&] {
      const size_t s = i_;
      while (i_ < t_.size() && t_[i_] >= '0' && t_[i_] <= '9') ++i_;
      if (i_ == s) error("bad number");
    };
    digits();
    if (i_ < t_.size() && t_[i_] == '.') {
      ++i_;
      digits();
    }
    if (i_ < t_.size() && (t_[i_] == 'e' || t_[i_] == 'E')) {
      ++i_;
      if (i_ < t_.size() && (t_[i_] == '+' || t_[i_] == '-')) ++i_;
      digits();
    }
    const std::string token(t_.substr(start, i_ - start));
    return std::strtod(token.c_str(), nullptr);
  }

  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += char(cp);
    } else if (cp < 0x800) {
      out += char(0xc0 | (cp >> 6));
      out += char(0x80 | (cp & 0x3f));
    } else if (cp < 0x10000) {
      out += char(0xe0 | (cp >> 12));
      out += char(0x80 | ((cp >> 6) & 0x3f));
      out += char(0x80 | (cp & 0x3f));
    } else {
      out += char(0xf0 | (cp >> 18));
      out += char(0x80 | ((cp >> 12) & 0x3f));
      out += char(0x80 | ((cp >> 6) & 0x3f));
      out += char(0x80 | (cp & 0x3f));
    }
  }

  uint32_t hex4() {
    if (i_ + 4 > t_.size()) error("short \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      const char h = t_[i_++];
      v <<= 4;
      if (h >= '0' && h <= '9') v |= uint32_t(h - '0');
      else if (h >= 'a' && h <= 'f') v |= uint32_t(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= uint32_t(h - 'A' + 10);
      else error("bad \\u escape");
    }
    return v;
  }

  std::string string() {
    ++i_;  // opening quote
    std::string out;
    while (true) {
      if (i_ >= t_.size()) error("unterminated string");
      const char c = t_[i_++];
      if (c == '"') return out;
      if (static_cast<unsigned char>(c) < 0x20) error("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (i_ >= t_.size()) error("unterminated escape");
      const char e = t_[i_++];
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
          uint32_t cp = hex4();
          if (cp >= 0xd800 && cp < 0xdc00 && t_.substr(i_, 2) == "\\u") {
            i_ += 2;
            const uint32_t lo = hex4();
            if (lo >= 0xdc00 && lo < 0xe000) cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
            else error("bad surrogate pair");
          }
          utf8(out, cp);
          break;
        }
        default: error("bad escape");
      }
    }
  }
};

}  // namespace

const JsonValue* JsonValue::find(const std::string& key) const {
  if (type != Type::Object) return nullptr;
  const JsonValue* found = nullptr;
  for (const auto& kv : object)
    if (kv.first == key) found = &kv.second;
  return found;
}

double JsonValue::number_or(const std::string& key, double fallback) const {
  const JsonValue* v = find(key);
  return v != nullptr && v->type == Type::Number ? v->number : fallback;
}

std::string JsonValue::string_or(const std::string& key, const std::string& fallback) const {
  const JsonValue* v = find(key);
  return v != nullptr && v->type == Type::String ? v->string : fallback;
}

JsonValue parse_json(std::string_view text) { return Parser(text).document(); }

}  // namespace gs
