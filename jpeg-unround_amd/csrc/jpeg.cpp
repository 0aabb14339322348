// The JPEG reader: a file's quantized DCT coefficients, quantization tables, sampling
// factors, ICC profile and EXIF orientation (what c/jpegio/src/jpegio.c:187-310 takes from
// libjpeg), and its component planes by the accurate integer inverse DCT (jpegio.c:363-
// 420). Written from ITU-T T.81: markers (B.1-B.2), Huffman tables and decoding (C, F.2.2),
// sequential (F) and progressive (G) scans with restart intervals; arithmetic coding and
// the lossless and hierarchical processes are refused. Where the file is corrupt, it
// goes on as libjpeg does where it can (extraneous bytes, a segment that ends early, a bad
// code, a missing EOI) and counts a warning, whose first text is kept.
#include <algorithm>
#include <array>
#include <cstring>
#include <string>
#include <vector>

#include "files.hpp"

namespace unround {

namespace {

// Natural order of the zig-zag positions (T.81 figure A.6).
constexpr std::array<uint8_t, 64> kNatural = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Failure {
  int32_t status;
  std::string message;
};

constexpr int32_t kDecode = 2, kUnsupported = 3, kLimit = 4;

[[noreturn]] void fail(int32_t status, const std::string& message) { throw Failure{status, message}; }

std::string hex(int value) {
  static const char digits[] = "0123456789abcdef";
  return std::string("0x") + digits[(value >> 4) & 15] + digits[value & 15];
}

// A Huffman table (T.81 C): the canonical codes of BITS and HUFFVAL, with a lookahead of
// 9 bits for the short codes.
struct Huffman {
  bool defined = false;
  std::array<uint8_t, 256> values{};
  std::array<int32_t, 18> maxcode{};  // the largest code of each length, -1 where none
  std::array<int32_t, 17> valptr{};
  std::array<int32_t, 17> mincode{};
  std::array<uint16_t, 512> fast{};  // (length << 8) | value, 0 where the code is longer

  void build(const uint8_t counts[16], const uint8_t* symbols, int total) {
    std::copy(symbols, symbols + total, values.begin());
    int code = 0, k = 0;
    fast.fill(0);
    for (int length = 1; length <= 16; ++length) {
      valptr[length] = k;
      mincode[length] = code;
      for (int i = 0; i < counts[length - 1]; ++i, ++k, ++code) {
        if (length <= 9) {
          const int shift = 9 - length;
          for (int fill = 0; fill < (1 << shift); ++fill)
            fast[(code << shift) | fill] = static_cast<uint16_t>((length << 8) | values[k]);
        }
      }
      // The next code still fits the length: no code is all ones (T.81 C, as libjpeg checks).
      if (code >= (1 << length)) fail(kDecode, "Bogus Huffman table definition");
      maxcode[length] = counts[length - 1] > 0 ? code - 1 : -1;
      code <<= 1;
    }
    maxcode[17] = 0x7FFFFFFF;
    defined = true;
  }
};

struct Component {
  int id = 0, h = 1, v = 1, slot = 0;
  int dc_table = 0, ac_table = 0;
  uint32_t width = 0, height = 0, blocks_across = 0, blocks_down = 0;  // the blocks that carry data
  uint32_t padded_across = 0, padded_down = 0;                          // to whole MCUs
  bool has_table = false;
  std::array<uint16_t, 64> table{};
  std::vector<int16_t> coefficients;  // padded_down x padded_across x 64, natural order
  int dc_predictor = 0;
};

class Decoder {
 public:
  Decoder(const uint8_t* data, size_t size, uint64_t max_pixels, int32_t max_scans, bool warnings_are_errors)
      : data_(data), size_(size), max_pixels_(max_pixels > 0 ? max_pixels : uint64_t{1} << 28),
        max_scans_(max_scans > 0 ? max_scans : 500), strict_(warnings_are_errors) {}

  void read(JpegImage& image) {
    if (size_ < 2 || data_[0] != 0xFF || data_[1] != 0xD8) fail(kDecode, "Not a JPEG file: starts with " +
                                                                       hex(size_ > 0 ? data_[0] : 0) + " " +
                                                                       hex(size_ > 1 ? data_[1] : 0));
    at_ = 2;
    bool frame = false, ended = false;
    while (!ended) {
      const int marker = next_marker();
      if (marker < 0) {
        if (!eof_warned_) warn("Premature end of JPEG file");
        break;
      }
      switch (marker) {
        case 0xC0:
        case 0xC1:
        case 0xC2:
          if (frame) fail(kDecode, "Invalid JPEG file structure: two SOF markers");
          start_of_frame(marker == 0xC2);
          frame = true;
          break;
        case 0xC3:
        case 0xC7:
        case 0xCB:
        case 0xCF:
          fail(kUnsupported, "lossless JPEG is not supported");
        case 0xC5:
        case 0xC6:
        case 0xCD:
        case 0xCE:
          fail(kUnsupported, "hierarchical JPEG is not supported");
        case 0xC9:
        case 0xCA:
          fail(kUnsupported, "arithmetic-coded JPEG is not supported by this reader");
        case 0xC4:
          define_huffman();
          break;
        case 0xCC:
          skip_segment();  // arithmetic conditioning, of no use without arithmetic coding
          break;
        case 0xDB:
          define_quantization();
          break;
        case 0xDD:
          define_restart();
          break;
        case 0xDA:
          if (!frame) fail(kDecode, "Invalid JPEG file structure: SOS before SOF");
          start_of_scan();
          break;
        case 0xD9:
          ended = true;
          break;
        case 0xE0:
        case 0xE1:
        case 0xE2:
        case 0xEE:
          application(marker);
          break;
        case 0xDC:
          skip_segment();  // DNL: the height is given by the frame here
          break;
        default:
          if (marker >= 0xD0 && marker <= 0xD7) {
            warn("Corrupt JPEG data: found marker " + hex(marker) + " outside a scan");
          } else if ((marker >= 0xE3 && marker <= 0xEF) || marker == 0xFE || (marker >= 0xF0 && marker <= 0xFD)) {
            skip_segment();
          } else if (marker == 0xD8) {
            fail(kDecode, "Invalid JPEG file structure: two SOI markers");
          } else {
            fail(kDecode, "Unsupported marker type " + hex(marker));
          }
      }
    }
    if (!frame) fail(kDecode, "the file has no frame");
    for (const Component& c : components_) {
      if (!c.has_table) fail(kDecode, "the file ends before the data of every component");
    }
    assemble(image);
  }

 private:
  // ---------------------------------------------------------------------------------
  // Markers and segments

  // The next marker after the current place, skipping fill bytes; extraneous bytes before
  // it are a warning. -1 at the end of the data.
  int next_marker() {
    size_t skipped = discarded_;
    discarded_ = 0;
    while (at_ < size_) {
      if (data_[at_] != 0xFF) {
        ++at_;
        ++skipped;
        continue;
      }
      while (at_ < size_ && data_[at_] == 0xFF) ++at_;
      if (at_ >= size_) break;
      const int code = data_[at_++];
      if (code == 0x00) {
        skipped += 2;
        continue;
      }
      if (skipped > 0) {
        warn("Corrupt JPEG data: " + std::to_string(skipped) + " extraneous bytes before marker " + hex(code));
      }
      return code;
    }
    return -1;
  }

  // A segment's body (after its length), checked against the end of the data.
  std::pair<const uint8_t*, size_t> segment() {
    if (size_ - at_ < 2) fail(kDecode, "Premature end of JPEG file");
    const size_t length = (static_cast<size_t>(data_[at_]) << 8) | data_[at_ + 1];
    if (length < 2 || size_ - at_ < length) fail(kDecode, "Premature end of JPEG file");
    const uint8_t* body = data_ + at_ + 2;
    at_ += length;
    return {body, length - 2};
  }

  void skip_segment() { (void)segment(); }

  void start_of_frame(bool progressive) {
    auto [body, length] = segment();
    if (length < 6) fail(kDecode, "Bogus marker length");
    progressive_ = progressive;
    const int precision = body[0];
    height_ = (static_cast<uint32_t>(body[1]) << 8) | body[2];
    width_ = (static_cast<uint32_t>(body[3]) << 8) | body[4];
    const int count = body[5];
    if (precision != 8) fail(kUnsupported, std::to_string(precision) + "-bit JPEG is not supported");
    if (height_ == 0 || width_ == 0 || count == 0) fail(kDecode, "Empty JPEG image (DNL not supported)");
    if (length != 6 + 3 * static_cast<size_t>(count)) fail(kDecode, "Bogus marker length");
    if (count > 4) fail(kDecode, "Too many color components: " + std::to_string(count) + ", max 4");
    for (int i = 0; i < count; ++i) {
      Component c;
      c.id = body[6 + 3 * i];
      c.h = body[7 + 3 * i] >> 4;
      c.v = body[7 + 3 * i] & 15;
      c.slot = body[8 + 3 * i];
      if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) fail(kDecode, "Bogus sampling factors");
      if (c.slot > 3) fail(kDecode, "Bogus quantization table number");
      components_.push_back(c);
    }
    if (count != 1 && count != 3) {
      fail(kUnsupported, "a JPEG of " + std::to_string(count) + " components is not supported");
    }
    const uint64_t pixels = static_cast<uint64_t>(width_) * height_;
    if (pixels > max_pixels_) {
      fail(kLimit, "the image has " + std::to_string(width_) + " x " + std::to_string(height_) +
                       " pixels, more than " + std::to_string(max_pixels_));
    }
    for (const Component& c : components_) {
      max_h_ = std::max(max_h_, c.h);
      max_v_ = std::max(max_v_, c.v);
    }
    mcus_across_ = (width_ + 8 * max_h_ - 1) / (8 * max_h_);
    mcus_down_ = (height_ + 8 * max_v_ - 1) / (8 * max_v_);
    for (Component& c : components_) {
      c.width = static_cast<uint32_t>((static_cast<uint64_t>(width_) * c.h + max_h_ - 1) / max_h_);
      c.height = static_cast<uint32_t>((static_cast<uint64_t>(height_) * c.v + max_v_ - 1) / max_v_);
      c.blocks_across = static_cast<uint32_t>((static_cast<uint64_t>(width_) * c.h + 8 * max_h_ - 1) / (8 * max_h_));
      c.blocks_down = static_cast<uint32_t>((static_cast<uint64_t>(height_) * c.v + 8 * max_v_ - 1) / (8 * max_v_));
      c.padded_across = mcus_across_ * c.h;
      c.padded_down = mcus_down_ * c.v;
      c.coefficients.assign(static_cast<size_t>(c.padded_across) * c.padded_down * 64, 0);
    }
  }

  void define_huffman() {
    auto [body, length] = segment();
    size_t k = 0;
    while (k < length) {
      if (length - k < 17) fail(kDecode, "Bogus marker length");
      const int kind = body[k] >> 4, slot = body[k] & 15;
      if (kind > 1 || slot > 3) fail(kDecode, "Bogus Huffman table definition");
      uint8_t counts[16];
      int total = 0;
      for (int i = 0; i < 16; ++i) {
        counts[i] = body[k + 1 + i];
        total += counts[i];
      }
      if (total > 256 || length - k - 17 < static_cast<size_t>(total)) fail(kDecode, "Bogus Huffman table definition");
      // A DC symbol is the size of a difference, at most 15.
      for (int i = 0; i < total && kind == 0; ++i)
        if (body[k + 17 + i] > 15) fail(kDecode, "Bogus Huffman table definition");
      (kind == 0 ? dc_tables_ : ac_tables_)[slot].build(counts, body + k + 17, total);
      k += 17 + static_cast<size_t>(total);
    }
  }

  void define_quantization() {
    auto [body, length] = segment();
    size_t k = 0;
    while (k < length) {
      const int precision = body[k] >> 4, slot = body[k] & 15;
      if (slot > 3) fail(kDecode, "Bogus quantization table number");
      const size_t bytes = precision == 0 ? 64 : 128;
      if (precision > 1 || length - k - 1 < bytes) fail(kDecode, "Bogus marker length");
      for (int i = 0; i < 64; ++i) {
        const uint16_t value = precision == 0 ? body[k + 1 + i]
                                              : static_cast<uint16_t>((body[k + 1 + 2 * i] << 8) | body[k + 2 + 2 * i]);
        tables_[slot][kNatural[i]] = value;
      }
      defined_tables_[slot] = true;
      k += 1 + bytes;
    }
  }

  void define_restart() {
    auto [body, length] = segment();
    if (length != 2) fail(kDecode, "Bogus marker length");
    restart_interval_ = (body[0] << 8) | body[1];
  }

  void application(int marker) {
    auto [body, length] = segment();
    if (marker == 0xE0 && length >= 5 && std::memcmp(body, "JFIF\0", 5) == 0) saw_jfif_ = true;
    if (marker == 0xEE && length >= 12 && std::memcmp(body, "Adobe", 5) == 0) {
      saw_adobe_ = true;
      adobe_transform_ = body[11];
    }
    if (marker == 0xE1 && orientation_ == 0) orientation_ = exif_orientation(body, length);
    if (marker == 0xE2 && length >= 14 && std::memcmp(body, "ICC_PROFILE\0", 12) == 0) {
      icc_chunks_.push_back({body[12], body[13], std::vector<uint8_t>(body + 14, body + length)});
    }
  }

  // The Orientation (0x0112) of the first IFD of an EXIF APP1 segment, 1 to 8, or 0
  // (every offset checked against the segment).
  static int32_t exif_orientation(const uint8_t* body, size_t length) {
    if (length < 14 || std::memcmp(body, "Exif\0\0", 6) != 0) return 0;
    const uint8_t* tiff = body + 6;
    const size_t size = length - 6;
    bool little;
    if (tiff[0] == 'I' && tiff[1] == 'I') {
      little = true;
    } else if (tiff[0] == 'M' && tiff[1] == 'M') {
      little = false;
    } else {
      return 0;
    }
    const auto u16 = [&](const uint8_t* p) {
      return little ? static_cast<uint32_t>(p[0] | (p[1] << 8)) : static_cast<uint32_t>((p[0] << 8) | p[1]);
    };
    const auto u32 = [&](const uint8_t* p) {
      return little ? (static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
                       (static_cast<uint32_t>(p[2]) << 16) | (static_cast<uint32_t>(p[3]) << 24))
                    : ((static_cast<uint32_t>(p[0]) << 24) | (static_cast<uint32_t>(p[1]) << 16) |
                       (static_cast<uint32_t>(p[2]) << 8) | static_cast<uint32_t>(p[3]));
    };
    if (u16(tiff + 2) != 42) return 0;
    const size_t ifd = u32(tiff + 4);
    if (ifd > size || size - ifd < 2) return 0;
    const size_t count = u16(tiff + ifd);
    for (size_t i = 0; i < count; ++i) {
      const size_t entry = ifd + 2 + 12 * i;
      if (entry > size || size - entry < 12) return 0;
      const uint8_t* e = tiff + entry;
      if (u16(e) != 0x0112) continue;
      if (u16(e + 2) != 3 || u32(e + 4) != 1) return 0;
      const uint32_t value = u16(e + 8);
      return value >= 1 && value <= 8 ? static_cast<int32_t>(value) : 0;
    }
    return 0;
  }

  // ---------------------------------------------------------------------------------
  // Scans

  // The color space as libjpeg guesses it from the markers before the first scan (T.81
  // leaves it to the application): grey for one component; for three, YCbCr after a JFIF
  // marker, RGB or YCbCr by an Adobe marker's transform, else RGB for the component
  // identifiers 'R', 'G', 'B' and YCbCr for any others.
  void settle_color_space() {
    if (components_.size() == 1) {
      color_space_ = 1;
    } else if (saw_jfif_) {
      color_space_ = 2;
    } else if (saw_adobe_) {
      if (adobe_transform_ > 1) warn("Unknown Adobe color transform code " + std::to_string(adobe_transform_));
      color_space_ = adobe_transform_ == 0 ? 3 : 2;
    } else {
      const bool rgb = components_[0].id == 'R' && components_[1].id == 'G' && components_[2].id == 'B';
      color_space_ = rgb ? 3 : 2;
    }
  }

  void start_of_scan() {
    if (scans_ == 0) settle_color_space();
    auto [body, length] = segment();
    if (length < 1) fail(kDecode, "Bogus marker length");
    const int count = body[0];
    if (count < 1 || count > 4 || length != 4 + 2 * static_cast<size_t>(count)) fail(kDecode, "Bogus marker length");
    scans_ += 1;
    if (scans_ > max_scans_) fail(kLimit, "the file has more than " + std::to_string(max_scans_) + " scans");
    std::vector<Component*> parts;
    for (int i = 0; i < count; ++i) {
      const int id = body[1 + 2 * i];
      Component* found = nullptr;
      for (Component& c : components_)
        if (c.id == id && std::find(parts.begin(), parts.end(), &c) == parts.end()) found = &c;
      if (found == nullptr) fail(kDecode, "Invalid component ID " + std::to_string(id) + " in SOS");
      found->dc_table = body[2 + 2 * i] >> 4;
      found->ac_table = body[2 + 2 * i] & 15;
      if (found->dc_table > 3 || found->ac_table > 3) fail(kDecode, "Bogus Huffman table number");
      parts.push_back(found);
    }
    const int ss = body[1 + 2 * count], se = body[2 + 2 * count];
    const int ah = body[3 + 2 * count] >> 4, al = body[3 + 2 * count] & 15;
    if (progressive_) {
      const bool bad = ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && count != 1) || ah > 13 || al > 13;
      if (bad) {
        fail(kDecode, "Invalid progressive parameters Ss=" + std::to_string(ss) + " Se=" + std::to_string(se) +
                          " Ah=" + std::to_string(ah) + " Al=" + std::to_string(al));
      }
    } else if (ss != 0 || se != 63 || ah != 0 || al != 0) {
      warn("Invalid SOS parameters for sequential JPEG");
    }
    int blocks_in_mcu = 0;
    for (Component* c : parts) {
      blocks_in_mcu += count == 1 ? 1 : c->h * c->v;
      // The component's table as its first scan begins (libjpeg keeps that copy).
      if (!c->has_table) {
        if (!defined_tables_[c->slot]) fail(kDecode, "Quantization table " + hex(c->slot) + " was not defined");
        c->table = tables_[c->slot];
        c->has_table = true;
      }
    }
    if (blocks_in_mcu > 10) fail(kDecode, "Sampling factors too large for interleaved scan");
    scan_ = {parts, ss, se, ah, al};
    for (Component* c : parts) {
      const bool dc = !progressive_ || ss == 0;
      const bool ac = !progressive_ || ss > 0;
      if (dc && (!progressive_ || ah == 0) && !dc_tables_[c->dc_table].defined)
        fail(kDecode, "Huffman table " + hex(c->dc_table) + " was not defined");
      if (ac && !ac_tables_[c->ac_table].defined) fail(kDecode, "Huffman table " + hex(c->ac_table) + " was not defined");
    }
    decode_scan();
  }

  struct Scan {
    std::vector<Component*> parts;
    int ss = 0, se = 63, ah = 0, al = 0;
  };

  // The bits of the entropy-coded segment: bytes up to a marker, 0xFF 0x00 the byte 0xFF;
  // at a marker or the end the bits are zeros, and a warning says so once a scan.
  uint32_t bits_ = 0;
  int bit_count_ = 0;
  bool at_marker_ = false, short_warned_ = false, eof_warned_ = false;

  void reset_bits() {
    bits_ = 0;
    bit_count_ = 0;
    at_marker_ = false;
  }

  // At the end of a scan or a restart interval: whole bytes read ahead into the bit buffer
  // and not used are counted with the bytes before the next marker, as libjpeg counts them.
  void discard_bits() {
    const int real = bit_count_ - std::min(bit_count_, past_end_);
    discarded_ += static_cast<size_t>(std::max(real, 0) / 8);
    reset_bits();
  }

  void fill() {
    while (bit_count_ <= 24) {
      int byte = 0;
      if (!at_marker_ && at_ < size_) {
        byte = data_[at_];
        if (byte == 0xFF) {
          size_t next = at_ + 1;
          while (next < size_ && data_[next] == 0xFF) ++next;
          if (next < size_ && data_[next] == 0x00) {
            at_ = next + 1;
          } else {
            at_marker_ = true;  // the marker stays for next_marker
            byte = 0;
          }
        } else {
          ++at_;
        }
      } else {
        if (!at_marker_ && at_ >= size_ && !eof_warned_) {
          // The data ends within a scan: libjpeg's source inserts an EOI and says so.
          eof_warned_ = true;
          warn("Premature end of JPEG file");
        }
        at_marker_ = true;
      }
      if (at_marker_ && byte == 0) {
        // Zeros past the segment's end; counted so that running out is seen.
        past_end_ += 8;
      }
      bits_ |= static_cast<uint32_t>(byte) << (24 - bit_count_);
      bit_count_ += 8;
    }
  }

  int past_end_ = 0;

  uint32_t get_bits(int n) {
    if (n == 0) return 0;
    if (bit_count_ < n) fill();
    const uint32_t value = bits_ >> (32 - n);
    bits_ <<= n;
    bit_count_ -= n;
    note_short();
    return value;
  }

  void note_short() {
    if (past_end_ > 0 && bit_count_ < past_end_ && !short_warned_) {
      short_warned_ = true;
      warn("Corrupt JPEG data: premature end of data segment");
    }
  }

  int decode_symbol(const Huffman& table) {
    if (bit_count_ < 16) fill();
    const uint16_t quick = table.fast[bits_ >> 23];
    if (quick != 0) {
      const int length = quick >> 8;
      bits_ <<= length;
      bit_count_ -= length;
      note_short();
      return quick & 0xFF;
    }
    int length = 10;
    int32_t code = static_cast<int32_t>(bits_ >> (32 - length));
    while (length <= 16 && code > table.maxcode[length]) {
      ++length;
      code = static_cast<int32_t>(bits_ >> (32 - length));
    }
    if (length > 16) {
      warn("Corrupt JPEG data: bad Huffman code");
      bits_ <<= 16;
      bit_count_ -= 16;
      return 0;  // as libjpeg: a zero, the rest of the block taken as empty
    }
    bits_ <<= length;
    bit_count_ -= length;
    note_short();
    return table.values[static_cast<size_t>(table.valptr[length] + code - table.mincode[length]) & 255];
  }

  static int extend(uint32_t value, int size) {
    return size == 0 ? 0 : (value < (1u << (size - 1)) ? static_cast<int>(value) - (1 << size) + 1
                                                         : static_cast<int>(value));
  }

  int receive_extend(int size) { return extend(get_bits(size), size); }

  int16_t* block_of(Component& c, uint32_t row, uint32_t column) {
    return c.coefficients.data() + (static_cast<size_t>(row) * c.padded_across + column) * 64;
  }

  void decode_block(Component& c, int16_t* block) {
    const Scan& s = scan_;
    if (!progressive_) {
      const int size = decode_symbol(dc_tables_[c.dc_table]);
      c.dc_predictor += receive_extend(size);
      block[0] = static_cast<int16_t>(c.dc_predictor);
      const Huffman& ac = ac_tables_[c.ac_table];
      for (int k = 1; k < 64;) {
        const int rs = decode_symbol(ac);
        const int r = rs >> 4, z = rs & 15;
        if (z != 0) {
          k += r;
          if (k > 63) break;
          block[kNatural[k]] = static_cast<int16_t>(receive_extend(z));
          ++k;
        } else if (r == 15) {
          k += 16;
        } else {
          break;
        }
      }
      return;
    }
    if (s.ss == 0) {
      // DC scans (G.1.2.1): first, the difference at Al; refinement, one bit.
      if (s.ah == 0) {
        const int size = decode_symbol(dc_tables_[c.dc_table]);
        c.dc_predictor += receive_extend(size);
        block[0] = static_cast<int16_t>(static_cast<uint32_t>(c.dc_predictor) << s.al);
      } else if (get_bits(1) != 0) {
        block[0] = static_cast<int16_t>(block[0] | (1 << s.al));
      }
      return;
    }
    const Huffman& ac = ac_tables_[c.ac_table];
    if (s.ah == 0) {
      // AC first (G.1.2.2).
      if (eob_run_ > 0) {
        --eob_run_;
        return;
      }
      for (int k = s.ss; k <= s.se;) {
        const int rs = decode_symbol(ac);
        const int r = rs >> 4, z = rs & 15;
        if (z != 0) {
          k += r;
          if (k > 63) break;
          block[kNatural[k]] = static_cast<int16_t>(static_cast<uint32_t>(receive_extend(z)) << s.al);
          ++k;
        } else if (r < 15) {
          eob_run_ = (1 << r) - 1;
          if (r > 0) eob_run_ += static_cast<int>(get_bits(r));
          break;
        } else {
          k += 16;
        }
      }
      return;
    }
    // AC refinement (G.1.2.3): a correction bit for every coefficient already nonzero
    // that the band passes over, new coefficients of +-1 at Al.
    const int one = 1 << s.al, minus = -(1 << s.al);
    int k = s.ss;
    if (eob_run_ == 0) {
      for (; k <= s.se; ++k) {
        const int rs = decode_symbol(ac);
        int r = rs >> 4;
        const int z = rs & 15;
        int value = 0;
        if (z != 0) {
          if (z != 1) warn("Corrupt JPEG data: bad Huffman code");
          value = get_bits(1) != 0 ? one : minus;
        } else if (r != 15) {
          eob_run_ = 1 << r;
          if (r > 0) eob_run_ += static_cast<int>(get_bits(r));
          break;
        }
        for (; k <= s.se; ++k) {
          int16_t& coefficient = block[kNatural[k]];
          if (coefficient != 0) {
            if (get_bits(1) != 0 && (coefficient & one) == 0)
              coefficient = static_cast<int16_t>(coefficient >= 0 ? coefficient + one : coefficient + minus);
          } else {
            if (r == 0) break;
            --r;
          }
        }
        if (value != 0 && k <= s.se) block[kNatural[k]] = static_cast<int16_t>(value);
      }
    }
    if (eob_run_ > 0) {
      // The rest of the band: correction bits only.
      for (; k <= s.se; ++k) {
        int16_t& coefficient = block[kNatural[k]];
        if (coefficient != 0 && get_bits(1) != 0 && (coefficient & one) == 0)
          coefficient = static_cast<int16_t>(coefficient >= 0 ? coefficient + one : coefficient + minus);
      }
      --eob_run_;
    }
  }

  int eob_run_ = 0;
  Scan scan_;

  // At the end of a restart interval: the RSTn marker, the predictors and the bit buffer
  // reset.
  void restart(int& expected) {
    // Whatever bits are left of the byte are padding.
    discard_bits();
    past_end_ = 0;
    const size_t before = at_;
    const int marker = next_marker();
    if (marker == 0xD0 + expected) {
      expected = (expected + 1) & 7;
    } else if (marker >= 0xD0 && marker <= 0xD7) {
      warn("Corrupt JPEG data: found marker " + hex(marker) + " instead of RST" + std::to_string(expected));
      expected = (marker - 0xD0 + 1) & 7;
    } else {
      warn("Corrupt JPEG data: found marker " + hex(marker < 0 ? 0 : marker) + " instead of RST" +
           std::to_string(expected));
      at_ = before;  // leave the marker for the file's loop; the scan's data has ended
      ended_scan_ = true;
    }
    for (Component& c : components_) c.dc_predictor = 0;
    eob_run_ = 0;
  }

  bool ended_scan_ = false;
  size_t discarded_ = 0;

  void decode_scan() {
    reset_bits();
    past_end_ = 0;
    short_warned_ = false;
    ended_scan_ = false;
    eob_run_ = 0;
    for (Component& c : components_) c.dc_predictor = 0;
    int expected = 0;
    uint32_t done = 0;
    const std::vector<Component*>& parts = scan_.parts;
    if (parts.size() == 1) {
      Component& c = *parts[0];
      // A non-interleaved scan: the component's blocks that carry data, one to an MCU.
      for (uint32_t row = 0; row < c.blocks_down && !ended_scan_; ++row) {
        for (uint32_t column = 0; column < c.blocks_across; ++column) {
          if (restart_interval_ > 0 && done > 0 && done % restart_interval_ == 0) {
            restart(expected);
            if (ended_scan_) break;
          }
          decode_block(c, block_of(c, row, column));
          ++done;
        }
      }
    } else {
      for (uint32_t mcu_row = 0; mcu_row < mcus_down_ && !ended_scan_; ++mcu_row) {
        for (uint32_t mcu_column = 0; mcu_column < mcus_across_; ++mcu_column) {
          if (restart_interval_ > 0 && done > 0 && done % restart_interval_ == 0) {
            restart(expected);
            if (ended_scan_) break;
          }
          for (Component* c : parts) {
            for (int v = 0; v < c->v; ++v)
              for (int h = 0; h < c->h; ++h)
                decode_block(*c, block_of(*c, mcu_row * c->v + v, mcu_column * c->h + h));
          }
          ++done;
        }
      }
    }
    // The bytes left of the segment up to its marker are skipped by the marker loop.
    discard_bits();
  }

  // ---------------------------------------------------------------------------------

  void warn(const std::string& text) {
    warnings_ += 1;
    if (warnings_ == 1) warning_ = text;
    if (strict_) fail(kDecode, text);
  }

  void assemble(JpegImage& image) {
    image.width = width_;
    image.height = height_;
    image.color_space = color_space_;
    image.max_h = max_h_;
    image.max_v = max_v_;
    image.progressive = progressive_;
    image.arithmetic = false;
    image.exif_orientation = orientation_;
    image.warnings = warnings_;
    image.warning = warning_;
    image.components.clear();
    for (Component& c : components_) {
      JpegComponent out;
      out.id = c.id;
      out.h = c.h;
      out.v = c.v;
      out.slot = c.slot;
      out.width = c.width;
      out.height = c.height;
      out.blocks_across = c.blocks_across;
      out.blocks_down = c.blocks_down;
      out.table = c.table;
      out.coefficients.resize(static_cast<size_t>(c.blocks_across) * c.blocks_down * 64);
      for (uint32_t row = 0; row < c.blocks_down; ++row) {
        std::memcpy(out.coefficients.data() + static_cast<size_t>(row) * c.blocks_across * 64,
                    c.coefficients.data() + static_cast<size_t>(row) * c.padded_across * 64,
                    sizeof(int16_t) * c.blocks_across * 64);
      }
      image.components.push_back(std::move(out));
    }
    image.icc_profile = icc_profile();
  }

  // The ICC profile of the APP2 chunks: every chunk of the same count, each sequence
  // number once, joined in order; none where they do not agree.
  std::vector<uint8_t> icc_profile() {
    if (icc_chunks_.empty()) return {};
    const int total = icc_chunks_[0].count;
    if (total == 0 || static_cast<int>(icc_chunks_.size()) != total) {
      warn("the ICC profile's APP2 markers do not agree; it is left out");
      return {};
    }
    std::vector<const IccChunk*> ordered(static_cast<size_t>(total), nullptr);
    for (const IccChunk& chunk : icc_chunks_) {
      if (chunk.count != total || chunk.sequence < 1 || chunk.sequence > total || ordered[chunk.sequence - 1]) {
        warn("the ICC profile's APP2 markers do not agree; it is left out");
        return {};
      }
      ordered[chunk.sequence - 1] = &chunk;
    }
    std::vector<uint8_t> profile;
    for (const IccChunk* chunk : ordered) profile.insert(profile.end(), chunk->data.begin(), chunk->data.end());
    return profile;
  }

  struct IccChunk {
    int sequence, count;
    std::vector<uint8_t> data;
  };

  const uint8_t* data_;
  size_t size_;
  size_t at_ = 0;
  uint64_t max_pixels_;
  int32_t max_scans_;
  bool strict_;
  bool progressive_ = false;
  uint32_t width_ = 0, height_ = 0;
  int max_h_ = 1, max_v_ = 1;
  uint32_t mcus_across_ = 0, mcus_down_ = 0;
  int color_space_ = 1;
  bool saw_jfif_ = false, saw_adobe_ = false;
  int adobe_transform_ = 1;
  int orientation_ = 0;
  int restart_interval_ = 0;
  int scans_ = 0;
  int warnings_ = 0;
  std::string warning_;
  std::vector<Component> components_;
  std::array<std::array<uint16_t, 64>, 4> tables_{};
  std::array<bool, 4> defined_tables_{};
  std::array<Huffman, 4> dc_tables_, ac_tables_;
  std::vector<IccChunk> icc_chunks_;
};

// ---------------------------------------------------------------------------------------
// The accurate integer inverse DCT of IJG's libjpeg (jidctint.c, after Loeffler, Ligtenberg
// and Moschytz): products by constants c of 13 fractional bits, FIX(c) = round(c 2^13),
// two passes with 2 extra bits between them, results level-shifted by 128 and limited to
// 0..255 through a table that wraps beyond +-512 as libjpeg's does.

constexpr int kConstBits = 13, kPass1Bits = 2;
constexpr int64_t fix(double c) { return static_cast<int64_t>(c * (1 << kConstBits) + 0.5); }
constexpr int64_t k0_298631336 = fix(0.298631336), k0_390180644 = fix(0.390180644), k0_541196100 = fix(0.541196100),
                  k0_765366865 = fix(0.765366865), k0_899976223 = fix(0.899976223), k1_175875602 = fix(1.175875602),
                  k1_501321110 = fix(1.501321110), k1_847759065 = fix(1.847759065), k1_961570560 = fix(1.961570560),
                  k2_053119869 = fix(2.053119869), k2_562915447 = fix(2.562915447), k3_072711026 = fix(3.072711026);

constexpr int64_t descale(int64_t x, int n) { return (x + (int64_t{1} << (n - 1))) >> n; }

uint8_t range_limited(int64_t value) {
  const int index = static_cast<int>(value & 1023);
  if (index < 128) return static_cast<uint8_t>(index + 128);
  if (index < 512) return 255;
  if (index < 896) return 0;
  return static_cast<uint8_t>(index - 896);
}

// One pass of the 1-D transform over 8 inputs d[0..7] (already dequantized or from the
// first pass): the even and the odd part, into the 8 sums before their descaling.
void rotate(const int64_t d[8], int64_t out[8]) {
  int64_t z2 = d[2], z3 = d[6];
  int64_t z1 = (z2 + z3) * k0_541196100;
  const int64_t tmp2_even = z1 + z3 * -k1_847759065;
  const int64_t tmp3_even = z1 + z2 * k0_765366865;
  const int64_t tmp0_even = (d[0] + d[4]) * (int64_t{1} << kConstBits);
  const int64_t tmp1_even = (d[0] - d[4]) * (int64_t{1} << kConstBits);
  const int64_t tmp10 = tmp0_even + tmp3_even, tmp13 = tmp0_even - tmp3_even;
  const int64_t tmp11 = tmp1_even + tmp2_even, tmp12 = tmp1_even - tmp2_even;
  int64_t tmp0 = d[7], tmp1 = d[5], tmp2 = d[3], tmp3 = d[1];
  z1 = tmp0 + tmp3;
  z2 = tmp1 + tmp2;
  z3 = tmp0 + tmp2;
  int64_t z4 = tmp1 + tmp3;
  const int64_t z5 = (z3 + z4) * k1_175875602;
  tmp0 *= k0_298631336;
  tmp1 *= k2_053119869;
  tmp2 *= k3_072711026;
  tmp3 *= k1_501321110;
  z1 *= -k0_899976223;
  z2 *= -k2_562915447;
  z3 *= -k1_961570560;
  z4 *= -k0_390180644;
  z3 += z5;
  z4 += z5;
  tmp0 += z1 + z3;
  tmp1 += z2 + z4;
  tmp2 += z2 + z3;
  tmp3 += z1 + z4;
  out[0] = tmp10 + tmp3;
  out[7] = tmp10 - tmp3;
  out[1] = tmp11 + tmp2;
  out[6] = tmp11 - tmp2;
  out[2] = tmp12 + tmp1;
  out[5] = tmp12 - tmp1;
  out[3] = tmp13 + tmp0;
  out[4] = tmp13 - tmp0;
}

void inverse_block(const int16_t* coefficients, const std::array<uint16_t, 64>& table, uint8_t* out, size_t stride) {
  int64_t workspace[64];
  for (int column = 0; column < 8; ++column) {
    int64_t d[8], sums[8];
    for (int row = 0; row < 8; ++row)
      d[row] = static_cast<int64_t>(coefficients[row * 8 + column]) * table[row * 8 + column];
    rotate(d, sums);
    for (int row = 0; row < 8; ++row) workspace[row * 8 + column] = descale(sums[row], kConstBits - kPass1Bits);
  }
  for (int row = 0; row < 8; ++row) {
    int64_t sums[8];
    rotate(workspace + row * 8, sums);
    for (int column = 0; column < 8; ++column)
      out[row * stride + column] = range_limited(descale(sums[column], kConstBits + kPass1Bits + 3));
  }
}

}  // namespace

bool is_jpeg(const uint8_t* data, size_t size) { return data != nullptr && size >= 2 && data[0] == 0xFF && data[1] == 0xD8; }

bool read_jpeg(const uint8_t* data, size_t size, uint64_t max_pixels, int32_t max_scans, bool warnings_are_errors,
               JpegImage& image, std::string& message, int32_t& status) {
  image = JpegImage{};
  try {
    Decoder decoder(data, size, max_pixels, max_scans, warnings_are_errors);
    decoder.read(image);
    status = 0;
    return true;
  } catch (const Failure& failure) {
    image = JpegImage{};
    status = failure.status;
    message = failure.message;
    return false;
  } catch (const std::bad_alloc&) {
    image = JpegImage{};
    status = 5;
    message = "out of memory";
    return false;
  }
}

void jpeg_planes(const JpegImage& image, std::vector<std::vector<uint8_t>>& planes) {
  planes.clear();
  for (const JpegComponent& c : image.components) {
    const size_t width = static_cast<size_t>(c.blocks_across) * 8, height = static_cast<size_t>(c.blocks_down) * 8;
    std::vector<uint8_t> plane(width * height);
    for (uint32_t row = 0; row < c.blocks_down; ++row)
      for (uint32_t column = 0; column < c.blocks_across; ++column)
        inverse_block(c.coefficients.data() + (static_cast<size_t>(row) * c.blocks_across + column) * 64, c.table,
                      plane.data() + row * 8 * width + column * 8, width);
    planes.push_back(std::move(plane));
  }
}

}  // namespace unround
