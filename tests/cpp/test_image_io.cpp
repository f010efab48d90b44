// Image files and JSON, without a GPU.  With no arguments: PNG write/read
// round trips (gray8, gray16, RGBA8), area resampling against hand-computed
// means, decoder errors, and the JSON reader (escapes, numbers, nesting,
// errors).  `test_image_io decode <file> <channels> <out>` writes the decoded
// image (u32 width, height, channels, then the bytes) for the Python tests,
// which check the decoders against files written by an independent encoder.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gs/image_io.h"
#include "gs/json.h"
#include "gs/splat_scene.h"
#include "test_util.h"

namespace {

void test_png_round_trips() {
  const uint32_t w = 37, h = 23;
  std::vector<uint8_t> g8(w * h), rgba(w * h * 4);
  std::vector<uint16_t> g16(w * h);
  for (uint32_t i = 0; i < w * h; ++i) {
    g8[i] = uint8_t(i * 7 + 3);
    g16[i] = uint16_t(i * 977 + 11);
    for (int c = 0; c < 4; ++c) rgba[i * 4 + c] = uint8_t(i * (c + 3) + c);
  }
  const gs::Image8 a = gs::decode_image(gs::encode_png_gray8(g8.data(), w, h), 1);
  EXPECT(a.width == w && a.height == h && a.data == g8, "gray8 round trip");
  const gs::Image16 b = gs::decode_image16(gs::encode_png_gray16(g16.data(), w, h));
  EXPECT(b.width == w && b.height == h && b.data == g16, "gray16 round trip");
  const gs::Image8 c = gs::decode_image(gs::encode_png_rgba8(rgba.data(), w, h), 4);
  EXPECT(c.data == rgba, "rgba8 round trip");
  // gray8 widened to 16 bits by 257; gray8 as RGBA: gray replicated, opaque.
  const gs::Image16 d = gs::decode_image16(gs::encode_png_gray8(g8.data(), w, h));
  EXPECT(d.data[5] == g8[5] * 257u, "gray8 as 16-bit");
  const gs::Image8 e = gs::decode_image(gs::encode_png_gray8(g8.data(), w, h), 4);
  EXPECT(e.data[20] == g8[5] && e.data[21] == g8[5] && e.data[23] == 255, "gray8 as RGBA");
}

void test_resize_area() {
  // 4x2 -> 2x1: means of 2x2 blocks.
  const uint8_t src[8] = {10, 20, 30, 40, 50, 60, 70, 80};
  const auto a = gs::resize_area(src, 4, 2, 1, 2, 1);
  EXPECT(a.size() == 2 && a[0] == 35 && a[1] == 55, "2x2 means %u %u", a[0], a[1]);
  // 3 -> 2 columns: output 0 covers pixel 0 and half of pixel 1 -> (10 + 0.5 * 40) / 1.5 = 20.
  const uint8_t row[3] = {10, 40, 100};
  const auto b = gs::resize_area(row, 3, 1, 1, 2, 1);
  EXPECT(b[0] == 20 && b[1] == 80, "3->2: %u %u", b[0], b[1]);
  const auto c = gs::resize_area(src, 2, 2, 2, 2, 2);  // 2x2 pixels of 2 channels: all 8 bytes
  EXPECT(c.size() == 8 && std::memcmp(c.data(), src, 8) == 0, "same size copies");
}

void test_decode_errors() {
  int thrown = 0;
  for (const std::vector<uint8_t>& bytes :
       {std::vector<uint8_t>{}, std::vector<uint8_t>{'n', 'o', 't', ' ', 'a', ' ', 'p', 'n', 'g'},
        std::vector<uint8_t>{0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a, 0, 0}, std::vector<uint8_t>{0xff, 0xd8, 0xff}}) {
    try {
      gs::decode_image(bytes, 4);
    } catch (const std::runtime_error&) {
      ++thrown;
    }
  }
  EXPECT(thrown == 4, "damaged files must throw (%d of 4 did)", thrown);
  // A truncated PNG: the image data ends early.
  std::vector<uint8_t> g(64 * 64, 7);
  std::vector<uint8_t> png = gs::encode_png_gray8(g.data(), 64, 64);
  png.resize(png.size() - 30);
  bool threw = false;
  try {
    gs::decode_image(png, 1);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  EXPECT(threw, "a truncated PNG must throw");
}

void test_json() {
  const gs::JsonValue v = gs::parse_json(
      " {\"a\": [1, -2.5e2, 0.125, true, false, null], \"s\": \"x\\\"\\\\\\/\\n\\u00e9\\ud83d\\ude00\", "
      "\"o\": {\"k\": {}}, \"a\": 3} ");
  EXPECT(v.type == gs::JsonValue::Type::Object && v.object.size() == 4, "object");
  EXPECT(v.number_or("a", 0) == 3.0, "repeated key: the last one");
  const gs::JsonValue& arr = v.object[0].second;
  EXPECT(arr.array.size() == 6 && arr.array[1].number == -250.0 && arr.array[2].number == 0.125 &&
             arr.array[3].boolean && arr.array[5].type == gs::JsonValue::Type::Null,
         "array values");
  EXPECT(v.string_or("s", "") == std::string("x\"\\/\n\xc3\xa9\xf0\x9f\x98\x80"), "string escapes");
  EXPECT(v.find("o") != nullptr && v.find("o")->find("k") != nullptr && v.find("missing") == nullptr, "nesting");
  int thrown = 0;
  for (const char* bad : {"", "{", "[1,]", "{\"a\" 1}", "\"\\x\"", "01x", "[1] 2", "{\"a\":tru}"}) {
    try {
      gs::parse_json(bad);
    } catch (const std::runtime_error&) {
      ++thrown;
    }
  }
  EXPECT(thrown == 8, "malformed JSON must throw (%d of 8 did)", thrown);
  // Nesting is held on the heap: 400 levels parse, past 512 is refused with an error.
  const std::string deep = std::string(400, '[') + "7" + std::string(400, ']');
  const gs::JsonValue* d = nullptr;
  gs::JsonValue dv = gs::parse_json(deep);
  d = &dv;
  for (int k = 0; k < 400 && d != nullptr; ++k) d = d->array.empty() ? nullptr : &d->array[0];
  EXPECT(d != nullptr && d->type == gs::JsonValue::Type::Number && d->number == 7.0, "400 nested arrays");
  bool too_deep = false;
  try {
    gs::parse_json(std::string(600, '[') + std::string(600, ']'));
  } catch (const std::runtime_error& e) {
    too_deep = std::string(e.what()).find("nesting") != std::string::npos;
  }
  EXPECT(too_deep, "600 nested arrays must be refused");
}

int decode_to_file(const char* in, int channels, const char* out) {
  try {
    const gs::Image8 img = gs::decode_image(gs::read_file(in), channels);
    std::vector<uint8_t> bytes(12);
    const uint32_t head[3] = {img.width, img.height, img.channels};
    std::memcpy(bytes.data(), head, 12);
    bytes.insert(bytes.end(), img.data.begin(), img.data.end());
    gs::write_file(out, bytes);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "decode: %s\n", e.what());
    return 3;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 5 && std::string(argv[1]) == "decode") return decode_to_file(argv[2], std::atoi(argv[3]), argv[4]);
  test_png_round_trips();
  test_resize_area();
  test_decode_errors();
  test_json();
  if (test::failures() == 0) std::printf("test_image_io: ok\n");
  return test::failures() == 0 ? 0 : 1;
}
