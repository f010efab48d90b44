// Image files of a dataset (the role of engine/src/image_io.cpp, png_write.cpp
// and image_write.cpp): PNG of every colour type and bit depth (Adam7
// included) and baseline / extended-Huffman JPEG are decoded here; PNG is
// written through zlib.  Progressive and 12-bit JPEG are rejected by name.
#pragma once

#include <cstdint>
#include <vector>

namespace gs {

struct Image8 {
  uint32_t width = 0, height = 0, channels = 0;
  std::vector<uint8_t> data;  // row-major, `channels` bytes per pixel
};

struct Image16 {
  uint32_t width = 0, height = 0;
  std::vector<uint16_t> data;  // one channel
};

// PNG or JPEG (by signature) as `channels` = 1 (luma), 3 (RGB) or 4 (RGBA;
// alpha 255 where the file has none).  16-bit samples round to 8 bits.
// Throws std::runtime_error on anything it cannot read.
Image8 decode_image(const std::vector<uint8_t>& bytes, int channels);
// One 16-bit channel (depth maps): a 16-bit PNG as stored, 8-bit images
// widened by 257.  Colour images give their luma.
Image16 decode_image16(const std::vector<uint8_t>& bytes);

// Area (box) resampling of an 8-bit image: each output pixel is the mean of
// the source area it covers, partial pixels weighted by their overlap.
std::vector<uint8_t> resize_area(const uint8_t* src, uint32_t width, uint32_t height, int channels, uint32_t out_width,
                                 uint32_t out_height);

// PNG files (zlib-compressed, no filtering), as the capture writes masks,
// depth and confidence (engine/include/gs/png_write.h).
std::vector<uint8_t> encode_png_gray8(const uint8_t* pixels, uint32_t width, uint32_t height);
std::vector<uint8_t> encode_png_gray16(const uint16_t* pixels, uint32_t width, uint32_t height);
std::vector<uint8_t> encode_png_rgba8(const uint8_t* pixels, uint32_t width, uint32_t height);

}  // namespace gs
