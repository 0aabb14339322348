// The PNG reader (PNG specification, 2nd edition, sections 5-9 and 11): chunks with their
// CRCs, IHDR, the IDAT stream inflated by zlib, the five filters undone, Adam7's passes
// put back in place, and the samples unpacked to one level each. What it keeps and refuses
// follows c/jpegio/src/png.c:412-506: grey, grey and alpha, RGB and RGB and alpha of every
// depth PNG allows; palette files refused (their colours were chosen, not rounded); a
// tRNS chunk noted in the warning and not taken; an iCCP profile kept.
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>

#include "files.hpp"

namespace unround {

namespace {

constexpr std::array<uint8_t, 8> kSignature = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};

uint32_t big_endian(const uint8_t* p) {
  return (static_cast<uint32_t>(p[0]) << 24) | (static_cast<uint32_t>(p[1]) << 16) |
         (static_cast<uint32_t>(p[2]) << 8) | static_cast<uint32_t>(p[3]);
}

struct ReadFailure {
  int32_t status;
  std::string message;
};

[[noreturn]] void fail(int32_t status, const std::string& message) { throw ReadFailure{status, message}; }

// The statuses of the file layer (the reference's unround_jpegio_status).
constexpr int32_t kUnsupported = 3, kLimit = 4, kMemory = 5, kBroken = 7;

// Inflates a zlib stream into exactly `wanted` bytes (fewer: the stream ends early; more
// are left, and a warning says so).
std::vector<uint8_t> inflated(const std::vector<uint8_t>& stream, size_t wanted, std::string& warning) {
  std::vector<uint8_t> out(wanted);
  z_stream z{};
  if (inflateInit(&z) != Z_OK) fail(kBroken, "zlib could not start");
  z.next_in = const_cast<Bytef*>(stream.data());
  size_t given = 0, made = 0;
  int result = Z_OK;
  // zlib counts in 32 bits: feed and take at most 1 GiB at a time.
  constexpr size_t kStep = size_t{1} << 30;
  while (made < wanted) {
    if (z.avail_in == 0 && given < stream.size()) {
      z.next_in = const_cast<Bytef*>(stream.data() + given);
      z.avail_in = static_cast<uInt>(std::min(kStep, stream.size() - given));
      given += z.avail_in;
    }
    z.next_out = out.data() + made;
    z.avail_out = static_cast<uInt>(std::min(kStep, wanted - made));
    const uInt before = z.avail_out;
    result = inflate(&z, Z_NO_FLUSH);
    made += before - z.avail_out;
    if (result == Z_STREAM_END) break;
    if (result != Z_OK && result != Z_BUF_ERROR) {
      inflateEnd(&z);
      fail(kBroken, std::string("the compressed data is broken: ") + (z.msg != nullptr ? z.msg : "zlib error"));
    }
    if (result == Z_BUF_ERROR && z.avail_in == 0 && given >= stream.size()) break;
  }
  if (made < wanted) {
    inflateEnd(&z);
    fail(kBroken, "the image data ends early");
  }
  if (result != Z_STREAM_END) {
    // Bytes beyond the picture's, or the end of the stream not reached within it.
    unsigned char extra = 0;
    z.next_out = &extra;
    z.avail_out = 1;
    const int more = inflate(&z, Z_NO_FLUSH);
    if (z.avail_out == 0 && warning.empty()) warning = "the image data holds more than the picture";
    (void)more;
  }
  inflateEnd(&z);
  return out;
}

// A stream inflated whole (an iCCP profile), at most `limit` bytes.
bool inflate_all(const uint8_t* data, size_t size, size_t limit, std::vector<uint8_t>& out) {
  z_stream z{};
  if (inflateInit(&z) != Z_OK) return false;
  z.next_in = const_cast<Bytef*>(data);
  z.avail_in = static_cast<uInt>(size);
  out.clear();
  std::array<uint8_t, 16384> room{};
  int result = Z_OK;
  while (result == Z_OK) {
    z.next_out = room.data();
    z.avail_out = static_cast<uInt>(room.size());
    result = inflate(&z, Z_NO_FLUSH);
    out.insert(out.end(), room.data(), room.data() + (room.size() - z.avail_out));
    if (out.size() > limit) result = Z_MEM_ERROR;
  }
  inflateEnd(&z);
  return result == Z_STREAM_END;
}

uint8_t paeth(uint8_t a, uint8_t b, uint8_t c) {
  const int p = static_cast<int>(a) + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

// Undoes the filter of one row in place, `previous` the row above (unfiltered) or null.
void unfilter(uint8_t kind, uint8_t* row, const uint8_t* previous, size_t bytes, size_t step) {
  switch (kind) {
    case 0:
      return;
    case 1:
      for (size_t i = step; i < bytes; ++i) row[i] = static_cast<uint8_t>(row[i] + row[i - step]);
      return;
    case 2:
      if (previous != nullptr)
        for (size_t i = 0; i < bytes; ++i) row[i] = static_cast<uint8_t>(row[i] + previous[i]);
      return;
    case 3:
      for (size_t i = 0; i < bytes; ++i) {
        const unsigned left = i >= step ? row[i - step] : 0u, up = previous != nullptr ? previous[i] : 0u;
        row[i] = static_cast<uint8_t>(row[i] + ((left + up) >> 1));
      }
      return;
    case 4:
      for (size_t i = 0; i < bytes; ++i) {
        const uint8_t left = i >= step ? row[i - step] : 0, up = previous != nullptr ? previous[i] : 0;
        const uint8_t corner = (i >= step && previous != nullptr) ? previous[i - step] : 0;
        row[i] = static_cast<uint8_t>(row[i] + paeth(left, up, corner));
      }
      return;
    default:
      fail(kBroken, "a row has an unknown filter type " + std::to_string(kind));
  }
}

// The passes of a picture: all of it, or Adam7's seven (x0, y0, dx, dy).
struct Pass {
  uint32_t x0, y0, dx, dy;
};
constexpr std::array<Pass, 7> kAdam7 = {{{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4},
                                          {1, 0, 2, 2}, {0, 1, 1, 2}}};

uint32_t pass_extent(uint32_t size, uint32_t first, uint32_t step) {
  return size > first ? (size - first + step - 1) / step : 0u;
}

// The bytes of a row of `width` pixels.
uint64_t row_bytes(uint64_t width, uint32_t channels, uint32_t bits) {
  return (width * channels * bits + 7) / 8;
}

void read_file(const uint8_t* data, size_t size, uint64_t max_pixels, PngImage& image) {
  if (data == nullptr || size < kSignature.size() || !is_png(data, size)) fail(kBroken, "the data is not a PNG file");
  size_t at = kSignature.size();
  bool header = false, ended = false;
  uint32_t color_type = 0;
  std::vector<uint8_t> stream;
  bool image_data = false;
  while (!ended) {
    if (size - at < 12) {
      if (image_data) break;  // no IEND: the picture is read (libpng's image read stops before it)
      fail(kBroken, "the file ends early");
    }
    const uint32_t length = big_endian(data + at);
    if (length > 0x7FFFFFFFu || size - at - 12 < length) fail(kBroken, "the file ends early");
    const uint8_t* type = data + at + 4;
    const uint8_t* body = type + 4;
    const uint32_t stored = big_endian(body + length);
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), type, 4 + length);
    std::string name(reinterpret_cast<const char*>(type), 4);
    const bool critical = (type[0] & 0x20) == 0;
    at += 12 + static_cast<size_t>(length);
    if (static_cast<uint32_t>(crc) != stored) {
      if (critical) fail(kBroken, "the CRC of the " + name + " chunk is wrong");
      if (image.warning.empty()) image.warning = "the " + name + " chunk's CRC is wrong; it is left out";
      continue;
    }
    if (!header && name != "IHDR") fail(kBroken, "the file does not begin with IHDR");
    if (name == "IHDR") {
      if (header) fail(kBroken, "a second IHDR chunk");
      if (length != 13) fail(kBroken, "an IHDR chunk of " + std::to_string(length) + " bytes");
      header = true;
      image.width = big_endian(body);
      image.height = big_endian(body + 4);
      image.bits = body[8];
      color_type = body[9];
      if (image.width == 0 || image.height == 0 || image.width > 0x7FFFFFFFu || image.height > 0x7FFFFFFFu)
        fail(kBroken, "the picture is empty or too large for PNG");
      if (body[10] != 0 || body[11] != 0) fail(kBroken, "an unknown compression or filter method");
      if (body[12] > 1) fail(kBroken, "an unknown interlace method");
      image.interlaced = body[12] == 1;
      const uint32_t b = image.bits;
      const bool low = b == 1 || b == 2 || b == 4;
      bool allowed = false;
      switch (color_type) {
        case 0:
          image.channels = 1;
          allowed = low || b == 8 || b == 16;
          break;
        case 2:
          image.channels = 3;
          allowed = b == 8 || b == 16;
          break;
        case 3:
          allowed = low || b == 8;
          if (allowed) fail(kUnsupported, "a palette PNG file is not taken: its colours were chosen, not rounded");
          break;
        case 4:
          image.channels = 2;
          allowed = b == 8 || b == 16;
          break;
        case 6:
          image.channels = 4;
          allowed = b == 8 || b == 16;
          break;
        default:
          break;
      }
      if (!allowed) fail(kBroken, "a colour type and depth PNG does not have");
      const uint64_t pixels = static_cast<uint64_t>(image.width) * image.height;
      if (pixels > (max_pixels > 0 ? max_pixels : uint64_t{1} << 28))
        fail(kLimit, "the picture has more pixels than the options allow");
    } else if (name == "IDAT") {
      stream.insert(stream.end(), body, body + length);
      image_data = true;
    } else if (name == "IEND") {
      ended = true;
    } else if (name == "tRNS") {
      if (image.warning.empty()) image.warning = "the transparent colour of the tRNS chunk is not taken";
    } else if (name == "iCCP") {
      // A name of 1 to 79 bytes, a NUL, the compression method 0, then the profile in zlib.
      const uint8_t* end = static_cast<const uint8_t*>(std::memchr(body, 0, std::min<size_t>(length, 80)));
      std::vector<uint8_t> profile;
      const bool named = end != nullptr && end > body && static_cast<size_t>(end - body) + 2 <= length;
      const bool read = named && end[1] == 0 &&
                        inflate_all(end + 2, length - static_cast<size_t>(end + 2 - body), size_t{1} << 27, profile);
      // A profile holds its own size in its first four bytes, and a header of 128 bytes
      // and a tag count.
      if (read && profile.size() >= 132 && big_endian(profile.data()) == profile.size()) {
        image.icc_profile = std::move(profile);
      } else if (image.warning.empty()) {
        image.warning = "the iCCP chunk's profile is broken; it is left out";
      }
    } else if (critical && name != "PLTE") {
      fail(kBroken, "an unknown critical chunk " + name);
    }
  }
  if (!header) fail(kBroken, "the file has no IHDR chunk");
  if (!image_data) fail(kBroken, "the file has no image data");
  // The filtered rows of every pass, each with its filter byte.
  const uint32_t W = image.width, H = image.height, C = image.channels, bits = image.bits;
  std::vector<Pass> passes;
  if (image.interlaced) {
    passes.assign(kAdam7.begin(), kAdam7.end());
  } else {
    passes.push_back({0, 0, 1, 1});
  }
  uint64_t total = 0;
  for (const Pass& p : passes) {
    const uint64_t w = pass_extent(W, p.x0, p.dx), h = pass_extent(H, p.y0, p.dy);
    if (w > 0 && h > 0) total += h * (1 + row_bytes(w, C, bits));
  }
  if (total > (uint64_t{1} << 40)) fail(kLimit, "the picture is too large");
  std::vector<uint8_t> raw = inflated(stream, static_cast<size_t>(total), image.warning);
  image.samples.assign(static_cast<size_t>(W) * H * C, 0);
  const size_t step = std::max<size_t>(1, C * bits / 8);
  size_t offset = 0;
  for (const Pass& p : passes) {
    const uint32_t w = pass_extent(W, p.x0, p.dx), h = pass_extent(H, p.y0, p.dy);
    if (w == 0 || h == 0) continue;
    const size_t bytes = static_cast<size_t>(row_bytes(w, C, bits));
    const uint8_t* previous = nullptr;
    for (uint32_t r = 0; r < h; ++r) {
      uint8_t* row = raw.data() + offset + 1;
      unfilter(raw[offset], row, previous, bytes, step);
      previous = row;
      offset += 1 + bytes;
      const size_t y = static_cast<size_t>(p.y0) + static_cast<size_t>(r) * p.dy;
      uint16_t* out = image.samples.data() + y * W * C;
      for (uint32_t k = 0; k < w; ++k) {
        const size_t x = static_cast<size_t>(p.x0) + static_cast<size_t>(k) * p.dx;
        for (uint32_t c = 0; c < C; ++c) {
          const size_t index = static_cast<size_t>(k) * C + c;  // the sample's place in the row
          uint16_t level;
          if (bits == 16) {
            level = static_cast<uint16_t>((row[2 * index] << 8) | row[2 * index + 1]);
          } else if (bits == 8) {
            level = row[index];
          } else {
            const size_t bit = index * bits;
            const unsigned shift = 8 - bits - static_cast<unsigned>(bit % 8);
            level = static_cast<uint16_t>((row[bit / 8] >> shift) & ((1u << bits) - 1u));
          }
          out[x * C + c] = level;
        }
      }
    }
  }
}

}  // namespace

bool is_png(const uint8_t* data, size_t size) {
  return data != nullptr && size >= kSignature.size() && std::memcmp(data, kSignature.data(), kSignature.size()) == 0;
}

bool read_png(const uint8_t* data, size_t size, uint64_t max_pixels, PngImage& image, std::string& message,
              int32_t& status) {
  image = PngImage{};
  try {
    read_file(data, size, max_pixels, image);
    status = 0;
    return true;
  } catch (const ReadFailure& failure) {
    image = PngImage{};
    status = failure.status;
    message = failure.message;
    return false;
  } catch (const std::bad_alloc&) {
    image = PngImage{};
    status = kMemory;
    message = "out of memory";
    return false;
  }
}

}  // namespace unround
