// The file layer's C interface (include/unround_jpegio.h): the reference's
// c/jpegio/include/unround/jpegio.h, on this library's own readers and writers. No C++
// exception crosses it.
#include "unround_jpegio.h"

#include <zlib.h>

#include <cstdlib>
#include <cstring>
#include <string>

#include "files.hpp"

namespace {

void copy_message(char* buffer, size_t size, const std::string& text) {
  if (buffer == nullptr || size == 0) return;
  const size_t end = std::min(text.size(), size - 1);
  std::memcpy(buffer, text.data(), end);
  buffer[end] = '\0';
}

// A copy in memory that the C caller frees with std::free (through this layer's frees).
template <typename T>
T* released(const T* data, size_t count) {
  if (count == 0) return nullptr;
  T* out = static_cast<T*>(std::malloc(sizeof(T) * count));
  if (out != nullptr) std::memcpy(out, data, sizeof(T) * count);
  return out;
}

}  // namespace

extern "C" {

int32_t unround_jpegio_abi_version(void) { return UNROUND_JPEGIO_ABI_VERSION; }

const char* unround_jpegio_libpng_version(void) {
  static const std::string text = std::string("unround PNG layer (MI355X edition), zlib ") + zlibVersion();
  return text.c_str();
}

unround_jpegio_status unround_jpegio_read_png(const uint8_t* data, size_t size, uint64_t max_pixels,
                                              unround_jpegio_png_image* image, char* message, size_t message_size) {
  if (image == nullptr || (message == nullptr && message_size > 0)) return UNROUND_JPEGIO_ERROR_ARGUMENT;
  *image = unround_jpegio_png_image{};
  if (data == nullptr) {
    copy_message(message, message_size, "the data is not a PNG file");
    return UNROUND_JPEGIO_ERROR_ARGUMENT;
  }
  try {
    unround::PngImage read;
    std::string text;
    int32_t status = UNROUND_JPEGIO_OK;
    if (!unround::read_png(data, size, max_pixels, read, text, status)) {
      copy_message(message, message_size, text);
      return status;
    }
    unround_jpegio_png_image made{};
    made.width = read.width;
    made.height = read.height;
    made.channels = static_cast<int32_t>(read.channels);
    made.bits = static_cast<int32_t>(read.bits);
    made.samples = released(read.samples.data(), read.samples.size());
    made.icc_profile = released(read.icc_profile.data(), read.icc_profile.size());
    made.icc_profile_size = made.icc_profile != nullptr ? read.icc_profile.size() : 0;
    if (made.samples == nullptr || (!read.icc_profile.empty() && made.icc_profile == nullptr)) {
      std::free(made.samples);
      std::free(made.icc_profile);
      copy_message(message, message_size, "out of memory");
      return UNROUND_JPEGIO_ERROR_MEMORY;
    }
    *image = made;
    copy_message(message, message_size, read.warning);
    return UNROUND_JPEGIO_OK;
  } catch (...) {
    copy_message(message, message_size, "out of memory");
    return UNROUND_JPEGIO_ERROR_MEMORY;
  }
}

void unround_jpegio_png_image_free(unround_jpegio_png_image* image) {
  if (image == nullptr) return;
  std::free(image->samples);
  std::free(image->icc_profile);
  *image = unround_jpegio_png_image{};
}

}  // extern "C"
