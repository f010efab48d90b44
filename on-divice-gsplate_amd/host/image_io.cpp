// Dataset image files (the role of engine/src/image_io.cpp, png_write.cpp):
// a PNG decoder (zlib inflate, the five row filters, Adam7, every colour type
// and bit depth, tRNS), a baseline / extended-Huffman JPEG decoder (Huffman
// with a 9-bit lookahead table, float separable IDCT, centred bilinear chroma
// upsampling, JFIF YCbCr or Adobe RGB), area resampling and PNG writers.
#include "gs/image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace gs {
namespace {

[[noreturn]] void fail(const std::string& what) { throw std::runtime_error(what); }

uint32_t be32(const uint8_t* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}
uint32_t be16(const uint8_t* p) { return (uint32_t(p[0]) << 8) | uint32_t(p[1]); }

// Decoded samples before the caller's channel layout: 1 gray, 2 gray+alpha,
// 3 RGB, 4 RGBA; every sample in [0, max], max 255 or 65535.
struct Raw {
  uint32_t width = 0, height = 0, channels = 0, max = 255;
  std::vector<uint16_t> data;
};

bool is_png(const std::vector<uint8_t>& b) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  return b.size() >= 8 && std::memcmp(b.data(), sig, 8) == 0;
}
bool is_jpeg(const std::vector<uint8_t>& b) { return b.size() >= 3 && b[0] == 0xff && b[1] == 0xd8 && b[2] == 0xff; }

// ---- PNG ----------------------------------------------------------------------

struct Pass {
  uint32_t x0, y0, dx, dy;
};
const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
const Pass kWhole = {0, 0, 1, 1};

uint32_t pass_extent(uint32_t n, uint32_t start, uint32_t step) { return n > start ? (n - start + step - 1) / step : 0; }

uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return uint8_t(a);
  return uint8_t(pb <= pc ? b : c);
}

// Undoes the row filters of `rows` rows of `row_bytes` (each led by its filter byte), in place.
void unfilter(uint8_t* data, size_t row_bytes, uint32_t rows, size_t bpp) {
  std::vector<uint8_t> zero(row_bytes, 0);
  const uint8_t* prev = zero.data();
  for (uint32_t r = 0; r < rows; ++r) {
    uint8_t* row = data + size_t(r) * (row_bytes + 1);
    uint8_t* cur = row + 1;
    switch (row[0]) {
      case 0: break;
      case 1:
        for (size_t i = bpp; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + cur[i - bpp]);
        break;
      case 2:
        for (size_t i = 0; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + prev[i]);
        break;
      case 3:
        for (size_t i = 0; i < row_bytes; ++i) {
          const int left = i >= bpp ? cur[i - bpp] : 0;
          cur[i] = uint8_t(cur[i] + ((left + prev[i]) >> 1));
        }
        break;
      case 4:
        for (size_t i = 0; i < row_bytes; ++i) {
          const int left = i >= bpp ? cur[i - bpp] : 0, up_left = i >= bpp ? prev[i - bpp] : 0;
          cur[i] = uint8_t(cur[i] + paeth(left, prev[i], up_left));
        }
        break;
      default: fail("PNG: unknown row filter " + std::to_string(row[0]));
    }
    prev = cur;
  }
}

Raw decode_png(const std::vector<uint8_t>& b) {
  size_t pos = 8;
  uint32_t w = 0, h = 0, depth = 0, colour = 0, interlace = 0;
  bool have_header = false, have_end = false;
  std::vector<uint8_t> idat, palette, trns;
  while (pos + 12 <= b.size()) {
    const uint32_t len = be32(&b[pos]);
    if (len > b.size() - pos - 12) fail("PNG: chunk runs past the end of the file");
    const uint8_t* type = &b[pos + 4];
    const uint8_t* data = &b[pos + 8];
    if (std::memcmp(type, "IHDR", 4) == 0) {
      if (len < 13) fail("PNG: short IHDR");
      w = be32(data), h = be32(data + 4), depth = data[8], colour = data[9], interlace = data[12];
      if (data[10] != 0 || data[11] != 0) fail("PNG: unknown compression or filter method");
      have_header = true;
    } else if (std::memcmp(type, "PLTE", 4) == 0) {
      palette.assign(data, data + len);
    } else if (std::memcmp(type, "tRNS", 4) == 0) {
      trns.assign(data, data + len);
    } else if (std::memcmp(type, "IDAT", 4) == 0) {
      idat.insert(idat.end(), data, data + len);
    } else if (std::memcmp(type, "IEND", 4) == 0) {
      have_end = true;
      break;
    }
    pos += size_t(len) + 12;
  }
  if (!have_header) fail("PNG: no IHDR");
  if (!have_end && idat.empty()) fail("PNG: no image data");
  if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) fail("PNG: unsupported size");
  if (interlace > 1) fail("PNG: unknown interlace method");
  uint32_t spp = 0;  // samples per pixel as stored
  switch (colour) {
    case 0: spp = 1; break;
    case 2: spp = 3; break;
    case 3: spp = 1; break;
    case 4: spp = 2; break;
    case 6: spp = 4; break;
    default: fail("PNG: unknown colour type");
  }
  const bool depth_ok = depth == 8 || depth == 16 || ((colour == 0 || colour == 3) && (depth == 1 || depth == 2 || depth == 4));
  if (!depth_ok || (colour == 3 && depth == 16)) fail("PNG: unsupported bit depth");
  if (colour == 3 && palette.size() < 3) fail("PNG: palette image without PLTE");
  const size_t bits_pp = size_t(spp) * depth;
  const size_t bpp = std::max<size_t>(1, bits_pp / 8);

  const Pass* passes = interlace ? kAdam7 : &kWhole;
  const int npasses = interlace ? 7 : 1;
  size_t expected = 0;
  for (int p = 0; p < npasses; ++p) {
    const uint32_t pw = pass_extent(w, passes[p].x0, passes[p].dx), ph = pass_extent(h, passes[p].y0, passes[p].dy);
    if (pw && ph) expected += size_t(ph) * (1 + (size_t(pw) * bits_pp + 7) / 8);
  }
  std::vector<uint8_t> raw(expected);
  {
    z_stream s{};
    if (inflateInit(&s) != Z_OK) fail("PNG: zlib init failed");
    s.next_in = idat.data();
    s.avail_in = uInt(idat.size());
    s.next_out = raw.data();
    s.avail_out = uInt(raw.size());
    const int rc = inflate(&s, Z_FINISH);
    const size_t got = raw.size() - s.avail_out;
    inflateEnd(&s);
    if ((rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) || got != expected) fail("PNG: image data is damaged");
  }

  // Stored samples -> Raw (palette expanded, tRNS keys turned into alpha, low depths scaled to 8 bits).
  const bool key_alpha = !trns.empty() && (colour == 0 || colour == 2);
  const bool pal_alpha = !trns.empty() && colour == 3;
  Raw out;
  out.width = w, out.height = h, out.max = depth == 16 ? 65535u : 255u;
  out.channels = colour == 3 ? (pal_alpha ? 4u : 3u) : (colour == 0 ? 1u : colour == 2 ? 3u : spp);
  if (key_alpha) out.channels += 1;
  out.data.assign(size_t(w) * h * out.channels, 0);
  const uint32_t low_scale = depth < 8 ? 255u / ((1u << depth) - 1u) : 1u;
  uint16_t key[3] = {0, 0, 0};
  if (key_alpha) {
    for (uint32_t c = 0; c < (colour == 0 ? 1u : 3u) && 2 * c + 1 < trns.size(); ++c) key[c] = uint16_t(be16(&trns[2 * c]));
  }
  size_t offset = 0;
  for (int p = 0; p < npasses; ++p) {
    const Pass& ps = passes[p];
    const uint32_t pw = pass_extent(w, ps.x0, ps.dx), ph = pass_extent(h, ps.y0, ps.dy);
    if (!pw || !ph) continue;
    const size_t row_bytes = (size_t(pw) * bits_pp + 7) / 8;
    unfilter(raw.data() + offset, row_bytes, ph, bpp);
    for (uint32_t r = 0; r < ph; ++r) {
      const uint8_t* row = raw.data() + offset + size_t(r) * (row_bytes + 1) + 1;
      for (uint32_t i = 0; i < pw; ++i) {
        uint32_t s[4] = {0, 0, 0, 0};
        for (uint32_t c = 0; c < spp; ++c) {
          const size_t sample = size_t(i) * spp + c;
          if (depth == 16) {
            s[c] = be16(row + sample * 2);
          } else if (depth == 8) {
            s[c] = row[sample];
          } else {
            const size_t bit = sample * depth;
            s[c] = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
          }
        }
        uint16_t* o = &out.data[(size_t(ps.y0 + r * ps.dy) * w + ps.x0 + i * ps.dx) * out.channels];
        if (colour == 3) {
          const uint32_t idx = s[0];
          if (size_t(idx) * 3 + 2 >= palette.size()) fail("PNG: palette index out of range");
          for (int c = 0; c < 3; ++c) o[c] = palette[idx * 3 + c];
          if (pal_alpha) o[3] = idx < trns.size() ? trns[idx] : 255;
          continue;
        }
        const bool is_key = key_alpha && (colour == 0 ? s[0] == key[0] : (s[0] == key[0] && s[1] == key[1] && s[2] == key[2]));
        for (uint32_t c = 0; c < spp; ++c) o[c] = uint16_t(s[c] * low_scale);
        if (key_alpha) o[spp] = is_key ? 0 : uint16_t(out.max);
      }
    }
    offset += size_t(ph) * (row_bytes + 1);
  }
  return out;
}

// ---- JPEG ---------------------------------------------------------------------

// Natural (row-major) index of the k-th coefficient in zigzag order.
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

constexpr int kLookBits = 9;

struct Huffman {
  bool present = false;
  uint8_t vals[256] = {};
  int mincode[17] = {}, maxcode[18] = {}, valptr[17] = {};
  uint8_t fast_len[1 << kLookBits] = {};  // 0: code longer than kLookBits
  uint8_t fast_val[1 << kLookBits] = {};

  void build(const uint8_t counts[16], const uint8_t* symbols, int total) {
    std::memcpy(vals, symbols, size_t(total));
    std::memset(fast_len, 0, sizeof(fast_len));
    int code = 0, k = 0;
    for (int len = 1; len <= 16; ++len) {
      valptr[len] = k;
      mincode[len] = code;
      for (int i = 0; i < counts[len - 1]; ++i, ++k, ++code) {
        // A canonical code of `len` bits is below 2^len; a table whose counts
        // assign more codes than that is not a prefix code (and would index
        // past the lookup tables).
        if (code >= (1 << len)) fail("JPEG: bad Huffman table (more codes than their lengths allow)");
        if (len <= kLookBits) {
          const int shift = kLookBits - len;
          for (int f = 0; f < (1 << shift); ++f) {
            fast_len[(code << shift) | f] = uint8_t(len);
            fast_val[(code << shift) | f] = symbols[k];
          }
        }
      }
      maxcode[len] = counts[len - 1] ? code - 1 : -1;
      code <<= 1;
    }
    maxcode[17] = 0x7fffffff;
    present = true;
  }
};

// Entropy-coded bytes, most significant bit first; stuffed 0xFF00 read as
// 0xFF; at a marker it feeds zeros and stops.
struct BitReader {
  const uint8_t* p = nullptr;
  const uint8_t* end = nullptr;
  uint32_t buf = 0;
  int nbits = 0;
  bool at_marker = false;

  void fill() {
    while (nbits <= 24) {
      uint32_t byte = 0;
      if (!at_marker && p < end) {
        if (*p == 0xff) {
          if (p + 1 < end && p[1] == 0x00) {
            byte = 0xff;
            p += 2;
          } else {
            at_marker = true;
          }
        } else {
          byte = *p++;
        }
      }
      buf |= byte << (24 - nbits);
      nbits += 8;
    }
  }
  uint32_t peek(int n) {
    fill();
    return buf >> (32 - n);
  }
  void skip(int n) {
    buf <<= n;
    nbits -= n;
  }
  uint32_t get(int n) {
    if (n == 0) return 0;
    const uint32_t v = peek(n);
    skip(n);
    return v;
  }
  // Drops the partial byte and steps over the next RSTn marker.
  void restart() {
    buf = 0, nbits = 0, at_marker = false;
    while (p + 1 < end && !(p[0] == 0xff && p[1] >= 0xd0 && p[1] <= 0xd7)) ++p;
    if (p + 1 >= end) fail("JPEG: missing restart marker");
    p += 2;
  }
  int decode(const Huffman& h) {
    const uint32_t look = peek(kLookBits);
    if (h.fast_len[look]) {
      skip(h.fast_len[look]);
      return h.fast_val[look];
    }
    const uint32_t bits16 = peek(16);
    for (int len = kLookBits + 1; len <= 16; ++len) {
      const int code = int(bits16 >> (16 - len));
      if (code <= h.maxcode[len]) {
        skip(len);
        return h.vals[h.valptr[len] + code - h.mincode[len]];
      }
    }
    fail("JPEG: bad Huffman code");
  }
  int receive_extend(int s) {
    if (s == 0) return 0;
    const int v = int(get(s));
    return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
  }
};

struct Component {
  int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
  int plane_w = 0, plane_h = 0;  // MCU-padded
  int valid_w = 0, valid_h = 0;  // ceil(W h / hmax), ceil(H v / vmax)
  int dc_pred = 0;
  std::vector<uint8_t> plane;
};

// A[x][u] = C(u)/2 cos((2x + 1) u pi / 16), C(0) = 1/sqrt(2): out = A F A^T.
struct IdctTable {
  float a[8][8];
  IdctTable() {
    for (int x = 0; x < 8; ++x) {
      for (int u = 0; u < 8; ++u) {
        const double c = u == 0 ? std::sqrt(0.5) : 1.0;
        a[x][u] = float(0.5 * c * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0));
      }
    }
  }
};

void idct_block(const int coef[64], uint8_t* out, int stride) {
  static const IdctTable t;
  float tmp[8][8];  // tmp[v][x] = sum_u F[v][u] a[x][u]
  for (int v = 0; v < 8; ++v) {
    for (int x = 0; x < 8; ++x) {
      float s = 0.0f;
      for (int u = 0; u < 8; ++u) s += float(coef[v * 8 + u]) * t.a[x][u];
      tmp[v][x] = s;
    }
  }
  for (int y = 0; y < 8; ++y) {
    for (int x = 0; x < 8; ++x) {
      float s = 0.0f;
      for (int v = 0; v < 8; ++v) s += t.a[y][v] * tmp[v][x];
      const float level = std::nearbyint(s + 128.0f);
      out[y * stride + x] = uint8_t(std::min(std::max(level, 0.0f), 255.0f));
    }
  }
}

Raw decode_jpeg(const std::vector<uint8_t>& b) {
  const uint8_t* p = b.data() + 2;
  const uint8_t* end = b.data() + b.size();
  uint16_t qt[4][64] = {};
  bool qt_present[4] = {false, false, false, false};
  Huffman dc[4], ac[4];
  std::vector<Component> comps;
  int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  int restart_interval = 0, adobe_transform = -1;
  bool frame = false, done = false, scanned = false;
  while (!done) {
    // Next marker (fill bytes 0xFF allowed).
    while (p < end && *p != 0xff) ++p;
    while (p < end && *p == 0xff) ++p;
    if (p >= end) break;
    const uint8_t marker = *p++;
    if (marker == 0xd9) break;  // EOI
    if (marker >= 0xd0 && marker <= 0xd7) continue;
    if (marker == 0x01) continue;
    if (end - p < 2) fail("JPEG: truncated segment");
    const uint32_t len = be16(p);
    if (len < 2 || size_t(end - p) < len) fail("JPEG: segment runs past the end of the file");
    const uint8_t* seg = p + 2;
    const uint8_t* seg_end = p + len;
    p += len;
    switch (marker) {
      case 0xc0:
      case 0xc1: {
        if (seg_end - seg < 6) fail("JPEG: short frame header");
        if (seg[0] != 8) fail("JPEG: only 8-bit samples are supported");
        height = int(be16(seg + 1)), width = int(be16(seg + 3));
        const int n = seg[5];
        if (width == 0 || height == 0) fail("JPEG: size missing from the frame header (DNL is not supported)");
        if ((n != 1 && n != 3) || seg_end - seg < 6 + 3 * n) fail("JPEG: only 1 or 3 components are supported");
        comps.resize(size_t(n));
        for (int i = 0; i < n; ++i) {
          Component& c = comps[size_t(i)];
          c.id = seg[6 + 3 * i];
          c.h = seg[7 + 3 * i] >> 4, c.v = seg[7 + 3 * i] & 15;
          c.tq = seg[8 + 3 * i] & 3;
          if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) fail("JPEG: bad sampling factors");
          hmax = std::max(hmax, c.h), vmax = std::max(vmax, c.v);
        }
        mcux = (width + 8 * hmax - 1) / (8 * hmax), mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comps) {
          c.plane_w = mcux * c.h * 8, c.plane_h = mcuy * c.v * 8;
          c.valid_w = (width * c.h + hmax - 1) / hmax, c.valid_h = (height * c.v + vmax - 1) / vmax;
          c.plane.assign(size_t(c.plane_w) * c.plane_h, 0);
        }
        frame = true;
        break;
      }
      case 0xc2:
      case 0xc6:
      case 0xca:
      case 0xce: fail("JPEG: progressive files are not supported");
      case 0xc3:
      case 0xc5:
      case 0xc7:
      case 0xc9:
      case 0xcb:
      case 0xcd:
      case 0xcf: fail("JPEG: lossless, hierarchical or arithmetic coding is not supported");
      case 0xc4: {
        const uint8_t* q = seg;
        while (q < seg_end) {
          if (seg_end - q < 17) fail("JPEG: short Huffman table");
          const int tc = q[0] >> 4, th = q[0] & 15;
          if (tc > 1 || th > 3) fail("JPEG: bad Huffman table id");
          int total = 0;
          for (int i = 0; i < 16; ++i) total += q[1 + i];
          if (total > 256 || seg_end - q < 17 + total) fail("JPEG: bad Huffman table");
          (tc == 0 ? dc[th] : ac[th]).build(q + 1, q + 17, total);
          q += 17 + total;
        }
        break;
      }
      case 0xdb: {
        const uint8_t* q = seg;
        while (q < seg_end) {
          const int pq = q[0] >> 4, tq = q[0] & 15;
          if (tq > 3 || pq > 1 || seg_end - q < 1 + 64 * (pq + 1)) fail("JPEG: bad quantization table");
          for (int k = 0; k < 64; ++k) qt[tq][k] = uint16_t(pq ? be16(q + 1 + 2 * k) : q[1 + k]);
          qt_present[tq] = true;
          q += 1 + 64 * (pq + 1);
        }
        break;
      }
      case 0xdd:
        if (seg_end - seg < 2) fail("JPEG: short restart interval");
        restart_interval = int(be16(seg));
        break;
      case 0xee:
        if (seg_end - seg >= 12 && std::memcmp(seg, "Adobe", 5) == 0) adobe_transform = seg[11];
        break;
      case 0xda: {
        if (!frame) fail("JPEG: scan before the frame header");
        const int ns = seg[0];
        if (ns < 1 || ns > 4 || seg_end - seg < 4 + 2 * ns) fail("JPEG: bad scan header");
        std::vector<Component*> sc;
        for (int i = 0; i < ns; ++i) {
          Component* c = nullptr;
          for (Component& k : comps)
            if (k.id == seg[1 + 2 * i]) c = &k;
          if (c == nullptr) fail("JPEG: scan names an unknown component");
          c->td = seg[2 + 2 * i] >> 4, c->ta = seg[2 + 2 * i] & 15;
          if (c->td > 3 || c->ta > 3 || !dc[c->td].present || !ac[c->ta].present) fail("JPEG: missing Huffman table");
          if (!qt_present[c->tq]) fail("JPEG: missing quantization table");
          c->dc_pred = 0;
          sc.push_back(c);
        }
        BitReader br;
        br.p = seg_end;
        br.end = end;
        int coef[64];
        auto block = [&](Component& c, int bx, int by) {
          std::fill(coef, coef + 64, 0);
          const int t = br.decode(dc[c.td]);
          if (t > 11) fail("JPEG: bad DC coefficient");
          c.dc_pred += br.receive_extend(t);
          const uint16_t* q = qt[c.tq];
          coef[0] = c.dc_pred * q[0];
          for (int k = 1; k < 64;) {
            const int rs = br.decode(ac[c.ta]);
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
              if (r != 15) break;
              k += 16;
              continue;
            }
            k += r;
            if (k > 63) fail("JPEG: coefficient index out of range");
            coef[kZigzag[k]] = br.receive_extend(s) * q[k];
            ++k;
          }
          idct_block(coef, &c.plane[size_t(by) * 8 * c.plane_w + size_t(bx) * 8], c.plane_w);
        };
        int mcu = 0;
        auto maybe_restart = [&]() {
          if (restart_interval > 0 && mcu > 0 && mcu % restart_interval == 0) {
            br.restart();
            for (Component* c : sc) c->dc_pred = 0;
          }
        };
        if (ns == 1) {
          Component& c = *sc[0];
          const int bw = (c.valid_w + 7) / 8, bh = (c.valid_h + 7) / 8;
          for (int by = 0; by < bh; ++by) {
            for (int bx = 0; bx < bw; ++bx, ++mcu) {
              maybe_restart();
              block(c, bx, by);
            }
          }
        } else {
          for (int my = 0; my < mcuy; ++my) {
            for (int mx = 0; mx < mcux; ++mx, ++mcu) {
              maybe_restart();
              for (Component* c : sc) {
                for (int v = 0; v < c->v; ++v)
                  for (int h = 0; h < c->h; ++h) block(*c, mx * c->h + h, my * c->v + v);
              }
            }
          }
        }
        // Continue after the scan: the next marker that is not a restart.
        p = br.p;
        while (p + 1 < end && !(p[0] == 0xff && p[1] != 0x00 && !(p[1] >= 0xd0 && p[1] <= 0xd7))) ++p;
        scanned = true;
        break;
      }
      default: break;  // APPn, COM and the rest carry nothing the image needs
    }
  }
  if (!frame || !scanned) fail("JPEG: no image data");

  Raw out;
  out.width = uint32_t(width), out.height = uint32_t(height), out.max = 255;
  out.channels = comps.size() == 1 ? 1u : 3u;
  out.data.assign(size_t(width) * height * out.channels, 0);
  // Each component sampled at the output pixel centre, bilinear between its own samples.
  std::vector<float> sample(comps.size());
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      for (size_t ci = 0; ci < comps.size(); ++ci) {
        const Component& c = comps[ci];
        if (c.h == hmax && c.v == vmax) {
          sample[ci] = c.plane[size_t(y) * c.plane_w + x];
          continue;
        }
        const float sx = (x + 0.5f) * float(c.h) / float(hmax) - 0.5f;
        const float sy = (y + 0.5f) * float(c.v) / float(vmax) - 0.5f;
        const int x0 = int(std::floor(sx)), y0 = int(std::floor(sy));
        const float fx = sx - float(x0), fy = sy - float(y0);
        const int xa = std::clamp(x0, 0, c.valid_w - 1), xb = std::clamp(x0 + 1, 0, c.valid_w - 1);
        const int ya = std::clamp(y0, 0, c.valid_h - 1), yb = std::clamp(y0 + 1, 0, c.valid_h - 1);
        const float top = c.plane[size_t(ya) * c.plane_w + xa] * (1 - fx) + c.plane[size_t(ya) * c.plane_w + xb] * fx;
        const float bot = c.plane[size_t(yb) * c.plane_w + xa] * (1 - fx) + c.plane[size_t(yb) * c.plane_w + xb] * fx;
        sample[ci] = top * (1 - fy) + bot * fy;
      }
      uint16_t* o = &out.data[(size_t(y) * width + x) * out.channels];
      if (out.channels == 1) {
        o[0] = uint16_t(std::lround(std::clamp(sample[0], 0.0f, 255.0f)));
        continue;
      }
      float r = sample[0], g = sample[1], bl = sample[2];
      if (adobe_transform != 0) {  // JFIF YCbCr (full range)
        const float Y = sample[0], cb = sample[1] - 128.0f, cr = sample[2] - 128.0f;
        r = Y + 1.402f * cr;
        g = Y - 0.344136f * cb - 0.714136f * cr;
        bl = Y + 1.772f * cb;
      }
      o[0] = uint16_t(std::lround(std::clamp(r, 0.0f, 255.0f)));
      o[1] = uint16_t(std::lround(std::clamp(g, 0.0f, 255.0f)));
      o[2] = uint16_t(std::lround(std::clamp(bl, 0.0f, 255.0f)));
    }
  }
  return out;
}

Raw decode_raw(const std::vector<uint8_t>& bytes) {
  if (is_png(bytes)) return decode_png(bytes);
  if (is_jpeg(bytes)) return decode_jpeg(bytes);
  fail(bytes.empty() ? "empty file" : "not a PNG or JPEG file");
}

// BT.601 luma of a sample triple, in the samples' own range.
uint32_t luma(uint32_t r, uint32_t g, uint32_t b) { return (299u * r + 587u * g + 114u * b + 500u) / 1000u; }

// ---- PNG writing ----------------------------------------------------------------

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  for (int s = 24; s >= 0; s -= 8) v.push_back(uint8_t(x >> s));
}

std::vector<uint8_t> encode_png(const uint8_t* rows, uint32_t width, uint32_t height, uint8_t depth, uint8_t colour,
                                size_t row_bytes) {
  std::vector<uint8_t> filtered;
  filtered.reserve((row_bytes + 1) * height);
  for (uint32_t y = 0; y < height; ++y) {
    filtered.push_back(0);
    filtered.insert(filtered.end(), rows + size_t(y) * row_bytes, rows + size_t(y + 1) * row_bytes);
  }
  uLongf zlen = compressBound(uLong(filtered.size()));
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, filtered.data(), uLong(filtered.size()), 6) != Z_OK) fail("PNG: compression failed");
  z.resize(zlen);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  auto chunk = [&](const char* type, const uint8_t* data, size_t n) {
    put_be32(out, uint32_t(n));
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_be32(out, uint32_t(crc32(0L, out.data() + start, uInt(n + 4))));
  };
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, width);
  put_be32(ihdr, height);
  ihdr.insert(ihdr.end(), {depth, colour, 0, 0, 0});
  chunk("IHDR", ihdr.data(), ihdr.size());
  chunk("IDAT", z.data(), z.size());
  chunk("IEND", nullptr, 0);
  return out;
}

}  // namespace

Image8 decode_image(const std::vector<uint8_t>& bytes, int channels) {
  if (channels != 1 && channels != 3 && channels != 4) throw std::invalid_argument("decode_image: channels 1, 3 or 4");
  const Raw r = decode_raw(bytes);
  Image8 img;
  img.width = r.width, img.height = r.height, img.channels = uint32_t(channels);
  img.data.resize(size_t(r.width) * r.height * channels);
  const auto to8 = [&](uint32_t v) { return uint8_t(r.max == 255 ? v : (v * 255u + 32767u) / 65535u); };
  const size_t n = size_t(r.width) * r.height;
  for (size_t i = 0; i < n; ++i) {
    const uint16_t* s = &r.data[i * r.channels];
    uint32_t rgb[3], a = r.max;
    if (r.channels <= 2) {
      rgb[0] = rgb[1] = rgb[2] = s[0];
      if (r.channels == 2) a = s[1];
    } else {
      rgb[0] = s[0], rgb[1] = s[1], rgb[2] = s[2];
      if (r.channels == 4) a = s[3];
    }
    uint8_t* o = &img.data[i * channels];
    if (channels == 1) {
      o[0] = to8(r.channels <= 2 ? rgb[0] : luma(rgb[0], rgb[1], rgb[2]));
    } else {
      for (int c = 0; c < 3; ++c) o[c] = to8(rgb[c]);
      if (channels == 4) o[3] = to8(a);
    }
  }
  return img;
}

Image16 decode_image16(const std::vector<uint8_t>& bytes) {
  const Raw r = decode_raw(bytes);
  Image16 img;
  img.width = r.width, img.height = r.height;
  img.data.resize(size_t(r.width) * r.height);
  const uint32_t widen = r.max == 255 ? 257u : 1u;
  for (size_t i = 0; i < img.data.size(); ++i) {
    const uint16_t* s = &r.data[i * r.channels];
    const uint32_t v = r.channels <= 2 ? s[0] : luma(s[0], s[1], s[2]);
    img.data[i] = uint16_t(v * widen);
  }
  return img;
}

std::vector<uint8_t> resize_area(const uint8_t* src, uint32_t width, uint32_t height, int channels, uint32_t out_width,
                                 uint32_t out_height) {
  if (width == 0 || height == 0 || out_width == 0 || out_height == 0 || channels <= 0) {
    throw std::invalid_argument("resize_area: empty image");
  }
  if (width == out_width && height == out_height) {
    return std::vector<uint8_t>(src, src + size_t(width) * height * channels);
  }
  struct Tap {
    uint32_t index;
    float weight;
  };
  // Source pixels under each output pixel and their share of its area.
  auto taps = [](uint32_t n, uint32_t out_n) {
    std::vector<std::vector<Tap>> t(out_n);
    const double scale = double(n) / out_n;
    for (uint32_t d = 0; d < out_n; ++d) {
      const double lo = d * scale, hi = (d + 1) * scale;
      for (uint32_t i = uint32_t(std::floor(lo)); i < std::min<double>(std::ceil(hi), n); ++i) {
        const double overlap = std::min(hi, double(i) + 1.0) - std::max(lo, double(i));
        if (overlap > 0.0) t[d].push_back({i, float(overlap / scale)});
      }
    }
    return t;
  };
  const auto tx = taps(width, out_width), ty = taps(height, out_height);
  std::vector<float> rows(size_t(height) * out_width * channels, 0.0f);
  for (uint32_t y = 0; y < height; ++y) {
    for (uint32_t x = 0; x < out_width; ++x) {
      float* o = &rows[(size_t(y) * out_width + x) * channels];
      for (const Tap& t : tx[x]) {
        const uint8_t* s = &src[(size_t(y) * width + t.index) * channels];
        for (int c = 0; c < channels; ++c) o[c] += t.weight * s[c];
      }
    }
  }
  std::vector<uint8_t> out(size_t(out_width) * out_height * channels);
  std::vector<float> acc(static_cast<size_t>(channels));
  for (uint32_t y = 0; y < out_height; ++y) {
    for (uint32_t x = 0; x < out_width; ++x) {
      std::fill(acc.begin(), acc.end(), 0.0f);
      for (const Tap& t : ty[y]) {
        const float* s = &rows[(size_t(t.index) * out_width + x) * channels];
        for (int c = 0; c < channels; ++c) acc[size_t(c)] += t.weight * s[c];
      }
      uint8_t* o = &out[(size_t(y) * out_width + x) * channels];
      for (int c = 0; c < channels; ++c) o[c] = uint8_t(std::lround(std::clamp(acc[size_t(c)], 0.0f, 255.0f)));
    }
  }
  return out;
}

std::vector<uint8_t> encode_png_gray8(const uint8_t* pixels, uint32_t width, uint32_t height) {
  return encode_png(pixels, width, height, 8, 0, width);
}

std::vector<uint8_t> encode_png_gray16(const uint16_t* pixels, uint32_t width, uint32_t height) {
  std::vector<uint8_t> be(size_t(width) * height * 2);
  for (size_t i = 0; i < size_t(width) * height; ++i) {
    be[2 * i] = uint8_t(pixels[i] >> 8);
    be[2 * i + 1] = uint8_t(pixels[i]);
  }
  return encode_png(be.data(), width, height, 16, 0, size_t(width) * 2);
}

std::vector<uint8_t> encode_png_rgba8(const uint8_t* pixels, uint32_t width, uint32_t height) {
  return encode_png(pixels, width, height, 8, 6, size_t(width) * 4);
}

}  // namespace gs
