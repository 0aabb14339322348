// Pictures out of the library (docs/math.md 10; rust/unround/src/output.rs,
// rust/unround/src/tiff.rs, rust/unround/src/orientation.rs, c/jpegio/src/png.c:131-360):
// samples of 8 and 16 bits by rounding or by the ordered dither of a blue-noise matrix,
// PNG files (written here on zlib), PNM and binary64 TIFF, the check of an ICC profile,
// and a picture turned upright as EXIF says.
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "files.hpp"

// The rounding of every product and difference here is the one written: exactness of
// the levels depends on it.
#pragma clang fp contract(off)

namespace unround {

namespace {

constexpr int kSide = 64;

// ---------------------------------------------------------------------------------------
// The blue-noise matrix by Ulichney's void-and-cluster method in integers (docs/math.md
// 10): points on the 64 x 64 torus, the energy of a place the sum over the points of
// round(2^40 exp(-d^2 / 4.5)), d the torus distance; the first place in row order where
// several share the least or greatest energy.

struct Pattern {
  const std::array<int64_t, kSide * kSide>* gaussian;
  std::array<bool, kSide * kSide> points{};
  std::array<int64_t, kSide * kSide> energy{};

  void change(int place, int64_t sign) {
    const int row = place / kSide, column = place % kSide;
    for (int i = 0; i < kSide; ++i) {
      const int from_row = ((i - row) % kSide + kSide) % kSide;
      for (int j = 0; j < kSide; ++j) {
        const int from_column = ((j - column) % kSide + kSide) % kSide;
        energy[i * kSide + j] += sign * (*gaussian)[from_row * kSide + from_column];
      }
    }
  }
  void put(int place) {
    points[place] = true;
    change(place, 1);
  }
  void take(int place) {
    points[place] = false;
    change(place, -1);
  }
  // The first point of greatest energy.
  int tightest() const {
    int best = -1;
    for (int k = 0; k < kSide * kSide; ++k)
      if (points[k] && (best < 0 || energy[k] > energy[best])) best = k;
    return best;
  }
  // The first place without a point of least energy.
  int emptiest() const {
    int best = -1;
    for (int k = 0; k < kSide * kSide; ++k)
      if (!points[k] && (best < 0 || energy[k] < energy[best])) best = k;
    return best;
  }
};

std::array<uint16_t, kSide * kSide> make_ranks() {
  std::array<int64_t, kSide * kSide> gaussian{};
  for (int row = 0; row < kSide; ++row) {
    for (int column = 0; column < kSide; ++column) {
      const int down = std::min(row, kSide - row), across = std::min(column, kSide - column);
      const long double power = -static_cast<long double>(down * down + across * across) / 4.5L;
      gaussian[row * kSide + column] = std::llround(std::ldexp(std::exp(power), 40));
    }
  }
  constexpr int kInitial = 410;
  Pattern initial{&gaussian};
  for (int k = 0; k < kInitial; ++k) initial.put(initial.emptiest());
  bool settled = false;
  for (int step = 0; step < kSide * kSide * 16 && !settled; ++step) {
    const int out = initial.tightest();
    initial.take(out);
    const int into = initial.emptiest();
    initial.put(into);
    settled = into == out;
  }
  std::array<uint16_t, kSide * kSide> ranks{};
  Pattern pattern = initial;
  for (int rank = kInitial - 1; rank >= 0; --rank) {
    const int place = pattern.tightest();
    pattern.take(place);
    ranks[place] = static_cast<uint16_t>(rank);
  }
  pattern = initial;
  for (int rank = kInitial; rank < kSide * kSide; ++rank) {
    const int place = pattern.emptiest();
    pattern.put(place);
    ranks[place] = static_cast<uint16_t>(rank);
  }
  return ranks;
}

// ---------------------------------------------------------------------------------------
// Levels (docs/math.md 10). y = scale x is held as its rounding and the exact error of it.

double rounded_level(double y, double error, double top) {
  const double whole = std::floor(y);
  const double level = (y - whole == 0.5 && error < 0.0) ? whole : std::round(y);
  return std::clamp(level, 0.0, top);
}

double dithered_level(double y, double error, double threshold, double top) {
  const double whole = std::floor(y);
  const double fraction = y - whole;
  const double limit = 1.0 - threshold;
  const double above = (fraction > limit || (fraction == limit && error >= 0.0)) ? 1.0 : 0.0;
  return std::clamp(whole + above, 0.0, top);
}

// The shifts of the matrix, rows and columns, of channels 0 to 3 dithered apart.
constexpr int kShifts[4][2] = {{0, 0}, {0, 32}, {32, 0}, {32, 32}};

template <typename Level>
void levels_of(const double* samples, size_t height, size_t width, size_t channels, double scale, double top,
               Dither dither, Level* out) {
  const size_t count = height * width * channels;
  if (dither == Dither::none) {
    for (size_t k = 0; k < count; ++k) {
      const double y = samples[k] * scale;
      out[k] = static_cast<Level>(rounded_level(y, std::fma(samples[k], scale, -y), top));
    }
    return;
  }
  const std::array<uint16_t, kSide * kSide>& ranks = blue_noise_ranks();
  for (size_t i = 0; i < height; ++i) {
    for (size_t j = 0; j < width; ++j) {
      for (size_t c = 0; c < channels; ++c) {
        const int down = dither == Dither::blue_noise ? kShifts[c][0] : 0;
        const int across = dither == Dither::blue_noise ? kShifts[c][1] : 0;
        const size_t rank = ranks[((i + down) % kSide) * kSide + (j + across) % kSide];
        const double threshold = (2.0 * static_cast<double>(rank) + 1.0) / 8192.0;
        const size_t k = (i * width + j) * channels + c;
        const double y = samples[k] * scale;
        out[k] = static_cast<Level>(dithered_level(y, std::fma(samples[k], scale, -y), threshold, top));
      }
    }
  }
}

bool picture_shape(const double* samples, size_t height, size_t width, size_t channels, std::string& message) {
  if (samples == nullptr || height == 0 || width == 0 || channels < 1 || channels > 4) {
    message = "a picture of height x width x 1 to 4 samples, not empty: " + std::to_string(height) + " x " +
              std::to_string(width) + " x " + std::to_string(channels);
    return false;
  }
  return true;
}

bool all_finite(const double* samples, size_t count, std::string& message) {
  for (size_t k = 0; k < count; ++k) {
    if (!std::isfinite(samples[k])) {
      message = "a sample to quantize is not finite: " + std::to_string(samples[k]);
      return false;
    }
  }
  return true;
}

void put_big_endian(std::vector<uint8_t>& out, uint32_t value) {
  for (int shift = 24; shift >= 0; shift -= 8) out.push_back(static_cast<uint8_t>(value >> shift));
}

template <typename T>
void put_little_endian(std::vector<uint8_t>& out, T value) {
  uint8_t bytes[sizeof(T)];
  std::memcpy(bytes, &value, sizeof(T));
  out.insert(out.end(), bytes, bytes + sizeof(T));
}

// ---------------------------------------------------------------------------------------
// PNG.

void put_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* body, size_t length) {
  put_big_endian(out, static_cast<uint32_t>(length));
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  if (length > 0) out.insert(out.end(), body, body + length);
  const uLong crc = crc32(crc32(0L, Z_NULL, 0), out.data() + start, static_cast<uInt>(4 + length));
  put_big_endian(out, static_cast<uint32_t>(crc));
}

// A zlib stream made as libpng 1.6 makes those of the files the reference writes
// (c/jpegio/src/png.c:250-276 leaves every zlib setting to libpng but the level), so that
// the bytes are libpng's for the same zlib: memory level 8; the window 2^15 bytes, or, for
// at most 16384 bytes of input, the least power of two (at least 2^9) that holds the input
// and the 262 bytes deflate looks ahead; the window the header names then trimmed to the
// least that holds the input itself, which is all inflate needs (FCHECK made again). The
// input goes to deflate in the pieces libpng gives it and the output into buffers of
// libpng's sizes: a stored stream's (level 0) blocks end where those do.
class LibpngStream {
 public:
  // `first` and `next`: the sizes of the first output buffer and of the others.
  LibpngStream(size_t input_size, int level, int strategy, size_t first, size_t next)
      : input_size_(input_size), next_(next) {
    int window_bits = 15;
    if (input_size <= kSmall) {
      while (window_bits > 9 && input_size + kLookAhead <= (size_t{1} << (window_bits - 1))) --window_bits;
    }
    ready_ = deflateInit2(&z_, level, Z_DEFLATED, window_bits, 8, strategy) == Z_OK;
    buffer_.resize(first);
    z_.next_out = buffer_.data();
    z_.avail_out = static_cast<uInt>(first);
  }
  ~LibpngStream() {
    if (ready_) deflateEnd(&z_);
  }
  LibpngStream(const LibpngStream&) = delete;
  LibpngStream& operator=(const LibpngStream&) = delete;

  // Deflates `size` bytes with no flush, handing every full buffer to `emit`.
  template <typename Emit>
  bool feed(const uint8_t* data, size_t size, Emit&& emit) {
    if (!ready_) return false;
    z_.next_in = const_cast<Bytef*>(data);
    z_.avail_in = static_cast<uInt>(size);
    for (;;) {
      const int result = deflate(&z_, Z_NO_FLUSH);
      if (result != Z_OK) return false;
      if (z_.avail_out == 0) hand_over(buffer_.size(), emit);
      if (z_.avail_in == 0) return true;
    }
  }

  // Ends the stream, handing over the rest, the last buffer as far as it is filled.
  template <typename Emit>
  bool finish(Emit&& emit) {
    if (!ready_) return false;
    z_.next_in = nullptr;
    z_.avail_in = 0;
    for (;;) {
      const int result = deflate(&z_, Z_FINISH);
      if (result == Z_STREAM_END) {
        hand_over(buffer_.size() - z_.avail_out, emit);
        return true;
      }
      if (result != Z_OK || z_.avail_out != 0) return false;
      hand_over(buffer_.size(), emit);
    }
  }

 private:
  static constexpr size_t kLookAhead = 262, kSmall = 16384;

  template <typename Emit>
  void hand_over(size_t filled, Emit&& emit) {
    // The header: CMF = CINFO << 4 | 8 (deflate), CINFO = log2(window) - 8; FLG's low five
    // bits make CMF * 256 + FLG a multiple of 31.
    if (!handed_ && input_size_ <= kSmall && filled >= 2 && (buffer_[0] & 0x0F) == 8 && (buffer_[0] >> 4) <= 7) {
      int info = buffer_[0] >> 4;
      while (info > 0 && input_size_ <= (size_t{1} << (info + 7))) --info;
      const unsigned cmf = static_cast<unsigned>(info << 4) | 8u;
      const unsigned flags = buffer_[1] & 0xE0u;
      buffer_[0] = static_cast<uint8_t>(cmf);
      buffer_[1] = static_cast<uint8_t>(flags + 31u - (cmf * 256u + flags) % 31u);
    }
    handed_ = true;
    if (filled > 0) emit(buffer_.data(), filled);
    buffer_.resize(next_);
    z_.next_out = buffer_.data();
    z_.avail_out = static_cast<uInt>(next_);
  }

  z_stream z_{};
  bool ready_ = false, handed_ = false;
  size_t input_size_, next_;
  std::vector<uint8_t> buffer_;
};

// A profile in zlib as libpng compresses text (the whole input at once, into a first
// buffer of 1024 bytes and then of 8192).
bool compressed_text(const uint8_t* data, size_t size, int level, std::vector<uint8_t>& out) {
  out.clear();
  LibpngStream stream(size, level, Z_DEFAULT_STRATEGY, 1024, 8192);
  const auto keep = [&out](const uint8_t* bytes, size_t count) { out.insert(out.end(), bytes, bytes + count); };
  return stream.feed(data, size, keep) && stream.finish(keep);
}

uint8_t paeth_of(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return static_cast<uint8_t>(a);
  return static_cast<uint8_t>(pb <= pc ? b : c);
}

// Row `row` filtered by `kind` into out (its filter byte first).
void filter_row(int kind, const uint8_t* row, const uint8_t* previous, size_t bytes, size_t step, uint8_t* out) {
  out[0] = static_cast<uint8_t>(kind);
  for (size_t i = 0; i < bytes; ++i) {
    const int left = i >= step ? row[i - step] : 0, up = previous != nullptr ? previous[i] : 0;
    const int corner = (i >= step && previous != nullptr) ? previous[i - step] : 0;
    int predicted = 0;
    switch (kind) {
      case 1:
        predicted = left;
        break;
      case 2:
        predicted = up;
        break;
      case 3:
        predicted = (left + up) >> 1;
        break;
      case 4:
        predicted = paeth_of(left, up, corner);
        break;
      default:
        break;
    }
    out[1 + i] = static_cast<uint8_t>(row[i] - predicted);
  }
}

}  // namespace

const std::array<uint16_t, 4096>& blue_noise_ranks() {
  static const std::array<uint16_t, 4096> ranks = make_ranks();
  return ranks;
}

bool quantize(const double* samples, size_t height, size_t width, size_t channels, bool sixteen, Dither dither,
              std::vector<uint16_t>& levels, std::string& message) {
  if (!picture_shape(samples, height, width, channels, message)) return false;
  const size_t count = height * width * channels;
  if (!all_finite(samples, count, message)) return false;
  levels.resize(count);
  levels_of(samples, height, width, channels, sixteen ? 257.0 : 1.0, sixteen ? 65535.0 : 255.0, dither,
            levels.data());
  return true;
}

bool check_icc(const uint8_t* profile, uint64_t size, int32_t channels, std::string& reason) {
  const auto word = [&](size_t at) {
    return (static_cast<uint32_t>(profile[at]) << 24) | (static_cast<uint32_t>(profile[at + 1]) << 16) |
           (static_cast<uint32_t>(profile[at + 2]) << 8) | static_cast<uint32_t>(profile[at + 3]);
  };
  const auto text = [&](size_t at) {
    std::string out(reinterpret_cast<const char*>(profile + at), 4);
    for (char& c : out)
      if (static_cast<unsigned char>(c) < 32 || static_cast<unsigned char>(c) > 126) c = '?';
    return out;
  };
  const auto signature = [](const char* four) {
    return (static_cast<uint32_t>(static_cast<uint8_t>(four[0])) << 24) |
           (static_cast<uint32_t>(static_cast<uint8_t>(four[1])) << 16) |
           (static_cast<uint32_t>(static_cast<uint8_t>(four[2])) << 8) | static_cast<uint32_t>(four[3]);
  };
  constexpr uint64_t kHeader = 132;
  if (profile == nullptr) {
    reason = "no profile is given";
  } else if (channels < 1 || channels > 4) {
    reason = "a picture here has 1 to 4 channels, not " + std::to_string(channels);
  } else if (size < kHeader) {
    reason = "it is " + std::to_string(size) + " bytes, fewer than the 132 of its header";
  } else if (size > 0x7FFFFFFFu) {
    reason = "it is " + std::to_string(size) + " bytes, more than a PNG chunk holds";
  } else if (word(0) != size) {
    reason = "its header says " + std::to_string(word(0)) + " bytes, and it is " + std::to_string(size);
  } else if (profile[8] > 3 && size % 4 != 0) {
    reason = "it is of version " + std::to_string(profile[8]) + " and " + std::to_string(size) +
             " bytes, not a multiple of 4";
  } else if (word(128) > (size - kHeader) / 12) {
    reason = "its " + std::to_string(word(128)) + " tags do not fit in it";
  } else if (word(64) >= 0xFFFFu) {
    reason = "its rendering intent is " + std::to_string(word(64));
  } else if (word(36) != signature("acsp")) {
    reason = "its signature is '" + text(36) + "', not 'acsp'";
  } else if (word(16) != signature(channels <= 2 ? "GRAY" : "RGB ")) {
    reason = "its data color space is '" + text(16) + "', not the picture's '" +
             std::string(channels <= 2 ? "GRAY" : "RGB ") + "'";
  } else if (word(12) == signature("abst") || word(12) == signature("link")) {
    reason = "it is of the class '" + text(12) + "', which does not go with a picture";
  } else if (word(20) != signature("XYZ ") && word(20) != signature("Lab ")) {
    reason = "its connection space is '" + text(20) + "', neither 'XYZ ' nor 'Lab '";
  } else {
    const uint64_t tags = word(128);
    for (uint64_t t = 0; t < tags; ++t) {
      const size_t at = static_cast<size_t>(kHeader + 12 * t);
      const uint64_t start = word(at + 4), length = word(at + 8);
      if (start > size || length > size - start) {
        reason = "its tag '" + text(at) + "' lies outside it";
        return false;
      }
    }
    reason.clear();
    return true;
  }
  return false;
}

bool png_bytes(const void* samples, uint32_t width, uint32_t height, int32_t channels, int32_t bits,
               int32_t compression, const uint8_t* icc, uint64_t icc_size, std::vector<uint8_t>& file,
               std::string& warning, std::string& message, int32_t& status) {
  status = 1;  // the file layer's ARGUMENT
  if (width == 0 || height == 0) {
    message = "the picture is empty";
  } else if (width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) {
    message = "the picture is larger than PNG allows";
  } else if (channels < 1 || channels > 4) {
    message = "a PNG picture here has 1 to 4 channels";
  } else if (bits != 8 && bits != 16) {
    message = "a PNG picture here has 8 or 16 bits a sample";
  } else if (compression < -1 || compression > 9) {
    message = "the compression is -1 or 0 to 9";
  } else if (samples == nullptr) {
    message = "no samples are given";
  } else if (icc == nullptr && icc_size > 0) {
    message = "no ICC profile is given for its size";
  } else {
    message.clear();
  }
  if (!message.empty()) return false;
  bool with_icc = false;
  warning.clear();
  if (icc != nullptr) {
    std::string reason;
    with_icc = check_icc(icc, icc_size, channels, reason);
    if (!with_icc) {
      warning = "the ICC profile is not written: " + reason;
    } else if (icc_size > 8000000u) {
      warning = "the ICC profile is written, but it is " + std::to_string(icc_size) +
                " bytes, and PNG readers by default leave out a profile of more than 8000000";
    }
  }
  const size_t step = static_cast<size_t>(channels) * static_cast<size_t>(bits / 8);
  const size_t row_bytes = static_cast<size_t>(width) * step;
  file.clear();
  static const uint8_t kSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  file.insert(file.end(), kSignature, kSignature + 8);
  static const uint8_t kColourTypes[4] = {0, 4, 2, 6};
  std::vector<uint8_t> header;
  put_big_endian(header, width);
  put_big_endian(header, height);
  header.push_back(static_cast<uint8_t>(bits));
  header.push_back(kColourTypes[channels - 1]);
  header.push_back(0);
  header.push_back(0);
  header.push_back(0);
  put_chunk(file, "IHDR", header.data(), header.size());
  status = 6;  // the file layer's ENCODE
  if (with_icc) {
    // The name, its NUL, the compression method 0, then the profile in zlib; stored
    // uncompressed (zlib's level 0) where the chunk would be shorter than 92 bytes, which
    // readers of the chunk's first 81 bytes and a zlib stream's least would refuse.
    static const char kName[] = "ICC profile";
    std::vector<uint8_t> body(kName, kName + sizeof(kName));  // with the NUL
    body.push_back(0);
    std::vector<uint8_t> packed;
    if (!compressed_text(icc, static_cast<size_t>(icc_size), compression, packed)) {
      message = "zlib could not compress the ICC profile";
      return false;
    }
    if (body.size() + packed.size() < 92 && !compressed_text(icc, static_cast<size_t>(icc_size), 0, packed)) {
      message = "zlib could not store the ICC profile";
      return false;
    }
    body.insert(body.end(), packed.begin(), packed.end());
    put_chunk(file, "iCCP", body.data(), body.size());
  }
  // The rows, each by the filter of the least sum of its bytes taken as signed (the
  // usual heuristic of PNG encoders, the first of equal sums), 16-bit samples most
  // significant byte first; each filtered row deflated as libpng deflates it, with no flush,
  // into IDAT chunks of 8192 bytes (its output buffer's), the last shorter.
  LibpngStream stream((1 + row_bytes) * static_cast<size_t>(height), compression, Z_FILTERED, 8192, 8192);
  const auto chunk = [&file](const uint8_t* bytes, size_t count) { put_chunk(file, "IDAT", bytes, count); };
  std::vector<uint8_t> row(row_bytes), previous(row_bytes), trial(1 + row_bytes), best(1 + row_bytes);
  for (size_t y = 0; y < height; ++y) {
    if (bits == 8) {
      std::memcpy(row.data(), static_cast<const uint8_t*>(samples) + y * row_bytes, row_bytes);
    } else {
      const uint16_t* words = static_cast<const uint16_t*>(samples) + y * (row_bytes / 2);
      for (size_t k = 0; k < row_bytes / 2; ++k) {
        row[2 * k] = static_cast<uint8_t>(words[k] >> 8);
        row[2 * k + 1] = static_cast<uint8_t>(words[k] & 0xFF);
      }
    }
    uint64_t best_sum = std::numeric_limits<uint64_t>::max();
    for (int kind = 0; kind < 5; ++kind) {
      filter_row(kind, row.data(), y > 0 ? previous.data() : nullptr, row_bytes, step, trial.data());
      uint64_t sum = 0;
      for (size_t i = 1; i <= row_bytes; ++i) sum += static_cast<uint64_t>(std::abs(static_cast<int8_t>(trial[i])));
      if (sum < best_sum) {
        best_sum = sum;
        std::swap(best, trial);
      }
    }
    if (!stream.feed(best.data(), best.size(), chunk)) {
      message = "zlib could not compress the picture";
      return false;
    }
    std::swap(row, previous);
  }
  if (!stream.finish(chunk)) {
    message = "zlib could not compress the picture";
    return false;
  }
  put_chunk(file, "IEND", nullptr, 0);
  status = 0;
  return true;
}

bool pnm_bytes(const double* samples, size_t height, size_t width, size_t channels, bool sixteen, Dither dither,
               std::vector<uint8_t>& file, std::string& message) {
  if (!picture_shape(samples, height, width, channels, message)) return false;
  if (channels == 2 || channels == 4) {
    message = "PNM holds no alpha: write PNG or TIFF";
    return false;
  }
  std::vector<uint16_t> levels;
  if (!quantize(samples, height, width, channels, sixteen, dither, levels, message)) return false;
  const std::string header = std::string(channels == 1 ? "P5" : "P6") + "\n" + std::to_string(width) + " " +
                             std::to_string(height) + "\n" + (sixteen ? "65535" : "255") + "\n";
  file.assign(header.begin(), header.end());
  file.reserve(file.size() + levels.size() * (sixteen ? 2 : 1));
  for (const uint16_t level : levels) {
    if (sixteen) file.push_back(static_cast<uint8_t>(level >> 8));
    file.push_back(static_cast<uint8_t>(level & 0xFF));
  }
  return true;
}

bool tiff_bytes(const double* samples, size_t height, size_t width, size_t channels, bool ycbcr, const uint8_t* icc,
                uint64_t icc_size, std::vector<uint8_t>& file, std::string& warning, std::string& message) {
  if (samples == nullptr || height == 0 || width == 0 || channels < 1 || channels > 4) {
    message = "a picture is height x width x 1 to 4 samples, and not empty: " + std::to_string(height) + " x " +
              std::to_string(width) + " x " + std::to_string(channels);
    return false;
  }
  if (ycbcr && channels != 3) {
    message = "Y, Cb and Cr are three samples to a pixel, not " + std::to_string(channels);
    return false;
  }
  warning.clear();
  bool with_icc = false;
  if (icc != nullptr) {
    std::string reason;
    with_icc = check_icc(icc, icc_size, static_cast<int32_t>(channels), reason);
    if (!with_icc) warning = "the ICC profile is not written: " + reason;
  }
  // TIFF 6.0, little-endian, one strip of chunky binary64 samples (SampleFormat 3); the
  // layout: the header, the samples, the values that do not fit in an entry (each on a
  // word boundary), then the IFD with its entries in ascending order of tags.
  constexpr uint16_t kShort = 3, kLong = 4, kRational = 5, kUndefined = 7;
  const uint64_t data_length = static_cast<uint64_t>(height) * width * channels * 8;
  const bool elsewhere = channels > 2;
  std::vector<uint8_t> bits, formats, resolution, coefficients, reference;
  if (elsewhere) {
    for (size_t c = 0; c < channels; ++c) put_little_endian<uint16_t>(bits, 64);
    for (size_t c = 0; c < channels; ++c) put_little_endian<uint16_t>(formats, 3);
  }
  for (const uint32_t v : {1u, 1u}) put_little_endian<uint32_t>(resolution, v);
  if (ycbcr) {
    for (const uint32_t v : {299u, 1000u, 587u, 1000u, 114u, 1000u}) put_little_endian<uint32_t>(coefficients, v);
    for (const uint32_t v : {0u, 1u, 255u, 1u, 128u, 1u, 255u, 1u, 128u, 1u, 255u, 1u})
      put_little_endian<uint32_t>(reference, v);
  }
  const uint64_t bits_offset = 8 + data_length;
  const uint64_t format_offset = bits_offset + bits.size();
  const uint64_t resolution_offset = format_offset + formats.size();
  const uint64_t coefficients_offset = resolution_offset + resolution.size();
  const uint64_t reference_offset = coefficients_offset + coefficients.size();
  uint64_t profile_offset = reference_offset + reference.size();
  profile_offset += profile_offset % 2;
  const uint64_t values_end = profile_offset + (with_icc ? icc_size : 0);
  const uint64_t ifd_offset = values_end + values_end % 2;
  struct Entry {
    uint16_t tag, kind;
    uint32_t count;
    std::vector<uint8_t> value;
  };
  const auto u32 = [](uint64_t v) {
    std::vector<uint8_t> out;
    put_little_endian<uint32_t>(out, static_cast<uint32_t>(v));
    return out;
  };
  const auto u16s = [](std::initializer_list<uint16_t> values) {
    std::vector<uint8_t> out;
    for (const uint16_t v : values) put_little_endian<uint16_t>(out, v);
    return out;
  };
  const auto repeated = [](uint16_t value, size_t count) {
    std::vector<uint8_t> out;
    for (size_t k = 0; k < count; ++k) put_little_endian<uint16_t>(out, value);
    return out;
  };
  const uint16_t photometric = ycbcr ? 6 : (channels >= 3 ? 2 : 1);
  const auto count = static_cast<uint16_t>(channels);
  std::vector<Entry> entries = {
      {256, kLong, 1, u32(width)},
      {257, kLong, 1, u32(height)},
      {258, kShort, count, elsewhere ? u32(bits_offset) : repeated(64, channels)},
      {259, kShort, 1, u16s({1})},
      {262, kShort, 1, u16s({photometric})},
      {273, kLong, 1, u32(8)},
      {277, kShort, 1, u16s({count})},
      {278, kLong, 1, u32(height)},
      {279, kLong, 1, u32(data_length)},
      {282, kRational, 1, u32(resolution_offset)},
      {283, kRational, 1, u32(resolution_offset)},
      {284, kShort, 1, u16s({1})},
      {296, kShort, 1, u16s({1})},
  };
  if (channels == 2 || channels == 4) entries.push_back({338, kShort, 1, u16s({2})});
  entries.push_back({339, kShort, count, elsewhere ? u32(format_offset) : repeated(3, channels)});
  if (ycbcr) {
    entries.push_back({529, kRational, 3, u32(coefficients_offset)});
    entries.push_back({530, kShort, 2, u16s({1, 1})});
    entries.push_back({531, kShort, 1, u16s({1})});
    entries.push_back({532, kRational, 6, u32(reference_offset)});
  }
  if (with_icc) entries.push_back({34675, kUndefined, static_cast<uint32_t>(icc_size), u32(profile_offset)});
  const uint64_t total = ifd_offset + 2 + 12 * entries.size() + 4;
  if (total > (uint64_t{1} << 32)) {
    message = "a picture of " + std::to_string(height) + " x " + std::to_string(width) + " x " +
              std::to_string(channels) + " binary64 samples does not fit a classic TIFF";
    return false;
  }
  file.clear();
  file.reserve(static_cast<size_t>(total));
  file.push_back('I');
  file.push_back('I');
  put_little_endian<uint16_t>(file, 42);
  put_little_endian<uint32_t>(file, static_cast<uint32_t>(ifd_offset));
  const size_t count_of_samples = height * width * channels;
  const size_t at = file.size();
  file.resize(at + count_of_samples * 8);
  for (size_t k = 0; k < count_of_samples; ++k) {
    uint64_t b;
    std::memcpy(&b, &samples[k], 8);
    for (int byte = 0; byte < 8; ++byte) file[at + 8 * k + byte] = static_cast<uint8_t>(b >> (8 * byte));
  }
  for (const auto* part : {&bits, &formats, &resolution, &coefficients, &reference})
    file.insert(file.end(), part->begin(), part->end());
  file.resize(static_cast<size_t>(profile_offset), 0);
  if (with_icc) file.insert(file.end(), icc, icc + icc_size);
  file.resize(static_cast<size_t>(ifd_offset), 0);
  put_little_endian<uint16_t>(file, static_cast<uint16_t>(entries.size()));
  for (const Entry& e : entries) {
    put_little_endian<uint16_t>(file, e.tag);
    put_little_endian<uint16_t>(file, e.kind);
    put_little_endian<uint32_t>(file, e.count);
    uint8_t held[4] = {0, 0, 0, 0};
    std::memcpy(held, e.value.data(), std::min<size_t>(4, e.value.size()));
    file.insert(file.end(), held, held + 4);
  }
  put_little_endian<uint32_t>(file, 0);
  return true;
}

bool orient(const double* samples, size_t height, size_t width, size_t channels, int32_t orientation, double* turned,
            size_t& turned_height, size_t& turned_width, std::string& message) {
  if (orientation < 0 || orientation > 8) {
    message = "an EXIF orientation is 1 to 8, or 0 for none, not " + std::to_string(orientation);
    return false;
  }
  if (channels == 0) {
    message = "a picture of " + std::to_string(height) + " x " + std::to_string(width) + " pixels of 0 samples";
    return false;
  }
  const bool transposes = orientation >= 5;
  turned_height = transposes ? width : height;
  turned_width = transposes ? height : width;
  const size_t row = width * channels;
  if (!transposes) {
    // The stored rows, each forwards or backwards, from the top or the bottom.
    const bool up = orientation == 3 || orientation == 4, back = orientation == 2 || orientation == 3;
    for (size_t r = 0; r < height; ++r) {
      const double* from = samples + (up ? height - 1 - r : r) * row;
      double* to = turned + r * row;
      if (!back) {
        std::memcpy(to, from, sizeof(double) * row);
      } else {
        for (size_t j = 0; j < width; ++j)
          std::memcpy(to + j * channels, from + (width - 1 - j) * channels, sizeof(double) * channels);
      }
    }
    return true;
  }
  // Pixel (r, c) of the result is stored at (c or h-1-c, r or w-1-r), in tiles of 64.
  const bool up = orientation == 6 || orientation == 7, back = orientation == 7 || orientation == 8;
  const size_t rows = width, columns = height;
  constexpr size_t kTile = 64;
  for (size_t r0 = 0; r0 < rows; r0 += kTile) {
    for (size_t c0 = 0; c0 < columns; c0 += kTile) {
      for (size_t r = r0; r < std::min(r0 + kTile, rows); ++r) {
        const size_t column = back ? width - 1 - r : r;
        for (size_t c = c0; c < std::min(c0 + kTile, columns); ++c) {
          const size_t stored = up ? height - 1 - c : c;
          std::memcpy(turned + (r * columns + c) * channels, samples + stored * row + column * channels,
                      sizeof(double) * channels);
        }
      }
    }
  }
  return true;
}

}  // namespace unround
