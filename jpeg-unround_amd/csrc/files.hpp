// Files at the library's edge: PNG files read into the levels of their samples, for the
// reconstruction of a picture of rounded samples (rust/unround/src/rounded.rs:71-100, whose
// reading is c/jpegio/src/png.c:412-506 on libpng); and the pictures the library gives
// out, as samples of 8 and 16 bits (rounded or dithered), PNG, PNM and binary64 TIFF files
// (rust/unround/src/output.rs, tiff.rs, orientation.rs). On zlib alone.
#pragma once

#include <array>
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

// A JPEG file's component (c/jpegio/include/unround/jpegio.h:69-91): its identifier,
// sampling factors and table slot, its samples and the blocks that carry them, the
// quantization table (natural order) its slot held when its first scan began, and its
// coefficients, blocks_down x blocks_across blocks of 64 in natural order (no dummy blocks).
struct JpegComponent {
  int id = 0, h = 1, v = 1, slot = 0;
  uint32_t width = 0, height = 0, blocks_across = 0, blocks_down = 0;
  std::array<uint16_t, 64> table{};
  std::vector<int16_t> coefficients;
};

// A JPEG file as the reconstruction reads it (c/jpegio/include/unround/jpegio.h:93-108).
struct JpegImage {
  uint32_t width = 0, height = 0;
  int color_space = 1;  // 1 grey, 2 YCbCr, 3 RGB
  int max_h = 1, max_v = 1;
  bool progressive = false, arithmetic = false;
  int exif_orientation = 0;
  int warnings = 0;
  std::string warning;
  std::vector<JpegComponent> components;
  std::vector<uint8_t> icc_profile;
};

// Whether the data begins with SOI.
bool is_jpeg(const uint8_t* data, size_t size);

// Reads a JPEG file's coefficients and metadata; false with the reason and a status of the
// file layer (2 data it cannot read, 3 a JPEG outside this reader's scope, 4 more pixels than
// max_pixels (2^28 where 0) or scans than max_scans (500 where 0), 5 out of memory). A
// corrupt-data warning fails the read where warnings_are_errors.
bool read_jpeg(const uint8_t* data, size_t size, uint64_t max_pixels, int32_t max_scans, bool warnings_are_errors,
               JpegImage& image, std::string& message, int32_t& status);

// The component planes of a JPEG file: each component's blocks by the accurate integer
// inverse DCT of libjpeg (jidctint.c), level-shifted and limited to 0-255, no smoothing, no
// upsampling: blocks_down * 8 rows of blocks_across * 8 samples.
void jpeg_planes(const JpegImage& image, std::vector<std::vector<uint8_t>>& planes);

// How samples become levels of 8 or 16 bits (docs/math.md 10): rounded, halves away from
// 0; or the ordered dither of the blue-noise matrix, the channels of a pixel by the matrix
// shifted apart, or together.
enum class Dither : int32_t { none = 0, blue_noise = 1, together = 2 };

// The 64 x 64 ranks of the blue-noise matrix, row by row, each of 0 to 4095 once: Ulichney's
// void-and-cluster method in integers, made at the first call.
const std::array<uint16_t, 4096>& blue_noise_ranks();

// The levels of height x width x channels (1 to 4) interleaved samples: of 8 bits each
// sample, of 16 bits 257 times it, taken exactly, made as the dither says and clamped.
// false with the reason for an empty picture or a sample that is not finite.
bool quantize(const double* samples, size_t height, size_t width, size_t channels, bool sixteen, Dither dither,
              std::vector<uint16_t>& levels, std::string& message);

// Whether an ICC profile goes with a picture of the channels (1, 2: grey; 3, 4: RGB), by
// what PNG readers check of an iCCP chunk's profile (c/jpegio/src/png.c:131-208); the
// reason where it does not.
bool check_icc(const uint8_t* profile, uint64_t size, int32_t channels, std::string& reason);

// A PNG file of 8- or 16-bit samples (uint8_t, or uint16_t in this machine's order), not
// interlaced, at zlib's level compression (-1: zlib's default), with the ICC profile where
// check_icc takes it (else a warning says why). false with the reason and a status of the
// file layer (1 an argument refused, 6 the encoding failed).
bool png_bytes(const void* samples, uint32_t width, uint32_t height, int32_t channels, int32_t bits,
               int32_t compression, const uint8_t* icc, uint64_t icc_size, std::vector<uint8_t>& file,
               std::string& warning, std::string& message, int32_t& status);

// A PNM file (P5 grey, P6 RGB) of 8 bits, or of 16 bits most significant byte first.
bool pnm_bytes(const double* samples, size_t height, size_t width, size_t channels, bool sixteen, Dither dither,
               std::vector<uint8_t>& file, std::string& message);

// A little-endian baseline TIFF of binary64 samples as they are (1 to 4 channels, alpha
// unassociated; with ycbcr, JFIF's Y, Cb and Cr), with the ICC profile where it goes.
bool tiff_bytes(const double* samples, size_t height, size_t width, size_t channels, bool ycbcr, const uint8_t* icc,
                uint64_t icc_size, std::vector<uint8_t>& file, std::string& warning, std::string& message);

// The picture turned and mirrored as EXIF's orientation (1 to 8, 0 for none) says to show
// it, into `turned` (as many samples), every sample moved bit for bit.
bool orient(const double* samples, size_t height, size_t width, size_t channels, int32_t orientation, double* turned,
            size_t& turned_height, size_t& turned_width, std::string& message);

}  // namespace unround
