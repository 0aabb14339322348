// Files at the library's edge: PNG files read into the levels of their samples, for the
// reconstruction of a picture of rounded samples (rust/unround/src/rounded.rs:71-100, whose
// reading is c/jpegio/src/png.c:412-506 on libpng); here on zlib alone.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace unround {

// A PNG file's picture as it stores it: rows from the top, each from the left, the samples
// of a pixel together (grey, grey and alpha, RGB, RGB and alpha), each the level the file
// holds (below 8 bits the level itself, not scaled).
struct PngImage {
  uint32_t width = 0, height = 0;
  uint32_t channels = 0;  // 1 to 4
  uint32_t bits = 0;      // 1, 2, 4, 8 or 16
  bool interlaced = false;
  std::vector<uint16_t> samples;  // height x width x channels
  std::vector<uint8_t> icc_profile;
  std::string warning;  // the first warning of the reading, or empty
};

// Whether the data begins with PNG's signature.
bool is_png(const uint8_t* data, size_t size);

// Reads a PNG file; false with the reason and a status of the file layer where it fails:
// 7 a file it cannot read, 3 a palette file, 4 more pixels than max_pixels (2^28 where 0),
// 5 out of memory.
bool read_png(const uint8_t* data, size_t size, uint64_t max_pixels, PngImage& image, std::string& message,
              int32_t& status);

}  // namespace unround
