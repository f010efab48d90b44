// JSON reader (RFC 8259) in two layers: a lexer that cuts the text into
// tokens (the six punctuation marks, strings with every escape decoded to
// UTF-8, numbers checked against the RFC grammar, the three literals) and an
// assembler that builds the tree with an explicit stack of open containers,
// so a deeply nested document costs heap, not the call stack.  Errors throw
// std::runtime_error naming the byte offset where the document went wrong.
#include "gs/json.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>

namespace gs {
namespace {

constexpr size_t kMaxDepth = 512;

enum class Tok { LBrace, RBrace, LBracket, RBracket, Colon, Comma, Str, Num, True, False, Null, End };

struct Token {
  Tok kind = Tok::End;
  size_t at = 0;     // byte offset of its first character
  std::string text;  // decoded string (Str)
  double num = 0.0;  // value (Num)
};

[[noreturn]] void fail_at(size_t at, const std::string& what) {
  throw std::runtime_error("JSON: " + what + " at byte " + std::to_string(at));
}

void append_utf8(std::string& out, uint32_t cp) {
  // 1 to 4 bytes: the lead byte carries the length in its top bits.
  if (cp < 0x80) {
    out.push_back(char(cp));
    return;
  }
  int extra = cp < 0x800 ? 1 : cp < 0x10000 ? 2 : 3;
  const uint8_t lead[4] = {0x00, 0xc0, 0xe0, 0xf0};
  out.push_back(char(lead[extra] | (cp >> (6 * extra))));
  while (extra-- > 0) out.push_back(char(0x80 | ((cp >> (6 * extra)) & 0x3f)));
}

class Lexer {
 public:
  explicit Lexer(std::string_view s) : s_(s) {}

  Token next() {
    while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\t' || s_[pos_] == '\n' || s_[pos_] == '\r')) ++pos_;
    Token t;
    t.at = pos_;
    if (pos_ == s_.size()) return t;
    const char c = s_[pos_];
    static const std::string_view marks = "{}[]:,";
    const size_t m = marks.find(c);
    if (m != std::string_view::npos) {
      ++pos_;
      t.kind = static_cast<Tok>(m);
      return t;
    }
    if (c == '"') {
      t.kind = Tok::Str;
      t.text = string_body();
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      t.kind = Tok::Num;
      t.num = number();
    } else if (literal("true")) {
      t.kind = Tok::True;
    } else if (literal("false")) {
      t.kind = Tok::False;
    } else if (literal("null")) {
      t.kind = Tok::Null;
    } else {
      fail_at(pos_, "unexpected character");
    }
    return t;
  }

 private:
  std::string_view s_;
  size_t pos_ = 0;

  bool literal(std::string_view w) {
    if (s_.compare(pos_, w.size(), w) != 0) return false;
    pos_ += w.size();
    return true;
  }

  bool digit_at(size_t i) const { return i < s_.size() && s_[i] >= '0' && s_[i] <= '9'; }

  // -? (0 | [1-9][0-9]*) (. [0-9]+)? ([eE] [+-]? [0-9]+)?
  double number() {
    const size_t from = pos_;
    size_t i = pos_;
    if (s_[i] == '-') ++i;
    if (!digit_at(i)) fail_at(i, "bad number");
    if (s_[i] == '0') {
      ++i;
    } else {
      while (digit_at(i)) ++i;
    }
    if (i < s_.size() && s_[i] == '.') {
      if (!digit_at(++i)) fail_at(i, "bad number");
      while (digit_at(i)) ++i;
    }
    if (i < s_.size() && (s_[i] == 'e' || s_[i] == 'E')) {
      ++i;
      if (i < s_.size() && (s_[i] == '+' || s_[i] == '-')) ++i;
      if (!digit_at(i)) fail_at(i, "bad number");
      while (digit_at(i)) ++i;
    }
    pos_ = i;
    return std::strtod(std::string(s_.substr(from, i - from)).c_str(), nullptr);
  }

  uint32_t code_unit() {  // the four hex digits of a \u escape
    if (pos_ + 4 > s_.size()) fail_at(pos_, "short \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k, ++pos_) {
      const char h = s_[pos_];
      const int d = h >= '0' && h <= '9' ? h - '0' : h >= 'a' && h <= 'f' ? h - 'a' + 10 : h >= 'A' && h <= 'F' ? h - 'A' + 10 : -1;
      if (d < 0) fail_at(pos_, "bad \\u escape");
      v = v * 16 + uint32_t(d);
    }
    return v;
  }

  std::string string_body() {
    std::string out;
    ++pos_;  // the opening quote
    for (;;) {
      if (pos_ >= s_.size()) fail_at(pos_, "unterminated string");
      const unsigned char c = static_cast<unsigned char>(s_[pos_++]);
      if (c == '"') return out;
      if (c < 0x20) fail_at(pos_ - 1, "control character in string");
      if (c != '\\') {
        out.push_back(char(c));
        continue;
      }
      if (pos_ >= s_.size()) fail_at(pos_, "unterminated escape");
      const char e = s_[pos_++];
      static const std::string_view from = "\"\\/bfnrt", to = "\"\\/\b\f\n\r\t";
      const size_t k = from.find(e);
      if (k != std::string_view::npos) {
        out.push_back(to[k]);
      } else if (e == 'u') {
        uint32_t cp = code_unit();
        const bool high = cp >= 0xd800 && cp < 0xdc00;
        if (high && s_.compare(pos_, 2, "\\u") == 0) {
          pos_ += 2;
          const uint32_t low = code_unit();
          if (low < 0xdc00 || low >= 0xe000) fail_at(pos_, "bad surrogate pair");
          cp = 0x10000 + ((cp - 0xd800) << 10) + (low - 0xdc00);
        }
        append_utf8(out, cp);
      } else {
        fail_at(pos_ - 1, "bad escape");
      }
    }
  }
};

// One open container on the assembler's stack.  Between tokens of the loop
// below it always waits for a member: the close, or (after the first member)
// a comma and the next one.
struct Open {
  JsonValue node;
  std::string key;    // objects: the key of the member being read
  bool first = true;  // no member yet
};

}  // namespace

JsonValue parse_json(std::string_view text) {
  Lexer lex(text);
  std::vector<Open> stack;
  JsonValue root;
  bool done = false;
  // A finished value goes into the innermost open container, or is the document.
  auto place = [&](JsonValue&& v) {
    if (stack.empty()) {
      root = std::move(v);
      done = true;
      return;
    }
    Open& o = stack.back();
    if (o.node.type == JsonValue::Type::Array) o.node.array.push_back(std::move(v));
    else o.node.object.emplace_back(std::move(o.key), std::move(v));
    o.first = false;
  };
  while (!done) {
    Token t = lex.next();
    if (!stack.empty()) {
      Open& o = stack.back();
      const bool is_object = o.node.type == JsonValue::Type::Object;
      const Tok close = is_object ? Tok::RBrace : Tok::RBracket;
      if (t.kind == close) {
        JsonValue v = std::move(o.node);
        stack.pop_back();
        place(std::move(v));
        continue;
      }
      if (!o.first) {
        if (t.kind != Tok::Comma) fail_at(t.at, is_object ? "expected ',' or '}'" : "expected ',' or ']'");
        t = lex.next();
        if (t.kind == close) fail_at(t.at, "trailing comma");
      }
      if (is_object) {
        if (t.kind != Tok::Str) fail_at(t.at, "expected a key");
        o.key = std::move(t.text);
        const Token colon = lex.next();
        if (colon.kind != Tok::Colon) fail_at(colon.at, "expected ':'");
        t = lex.next();
      }
    }
    JsonValue v;
    switch (t.kind) {
      case Tok::LBrace:
      case Tok::LBracket: {
        if (stack.size() >= kMaxDepth) fail_at(t.at, "nesting too deep");
        Open o;
        o.node.type = t.kind == Tok::LBrace ? JsonValue::Type::Object : JsonValue::Type::Array;
        stack.push_back(std::move(o));
        continue;
      }
      case Tok::Str:
        v.type = JsonValue::Type::String;
        v.string = std::move(t.text);
        break;
      case Tok::Num:
        v.type = JsonValue::Type::Number;
        v.number = t.num;
        break;
      case Tok::True:
      case Tok::False:
        v.type = JsonValue::Type::Bool;
        v.boolean = t.kind == Tok::True;
        break;
      case Tok::Null:
        break;
      case Tok::End:
        fail_at(t.at, "unexpected end");
      default:
        fail_at(t.at, "expected a value");
    }
    place(std::move(v));
  }
  const Token rest = lex.next();
  if (rest.kind != Tok::End) fail_at(rest.at, "trailing characters");
  return root;
}

namespace {

// The last member named `key` whose type is `type`, else nullptr.
const JsonValue* member_of(const JsonValue& obj, const std::string& key) {
  if (obj.type != JsonValue::Type::Object) return nullptr;
  const auto hit = std::find_if(obj.object.rbegin(), obj.object.rend(), [&](const auto& kv) { return kv.first == key; });
  return hit == obj.object.rend() ? nullptr : &hit->second;
}

}  // namespace

const JsonValue* JsonValue::find(const std::string& key) const { return member_of(*this, key); }

double JsonValue::number_or(const std::string& key, double fallback) const {
  const JsonValue* m = member_of(*this, key);
  if (m == nullptr || m->type != Type::Number) return fallback;
  return m->number;
}

std::string JsonValue::string_or(const std::string& key, const std::string& fallback) const {
  const JsonValue* m = member_of(*this, key);
  if (m == nullptr || m->type != Type::String) return fallback;
  return m->string;
}

}  // namespace gs
