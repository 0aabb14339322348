// The file layer's C interface (include/unround_jpegio.h): the reference's
// c/jpegio/include/unround/jpegio.h, on this library's own readers and writers. No C++
// exception crosses it.
#include "unround_jpegio.h"
#include "unround.h"

#include <zlib.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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

const char* unround_jpegio_libjpeg_version(void) { return "unround JPEG reader (MI355X edition; ITU-T T.81)"; }

unround_jpegio_status unround_jpegio_read(const uint8_t* data, size_t size, const unround_jpegio_options* options,
                                          unround_jpegio_image* image, char* message, size_t message_size) {
  if (image == nullptr || (message == nullptr && message_size > 0)) return UNROUND_JPEGIO_ERROR_ARGUMENT;
  *image = unround_jpegio_image{};
  if (data == nullptr && size > 0) {
    copy_message(message, message_size, "the data is not addressable as one buffer");
    return UNROUND_JPEGIO_ERROR_ARGUMENT;
  }
  try {
    unround::JpegImage read;
    std::string text;
    int32_t status = UNROUND_JPEGIO_OK;
    const bool done = unround::read_jpeg(data, size, options != nullptr ? options->max_pixels : 0,
                                         options != nullptr ? options->max_scans : 0,
                                         options != nullptr && options->warnings_are_errors != 0, read, text, status);
    if (!done) {
      copy_message(message, message_size, text);
      return status;
    }
    unround_jpegio_image made{};
    made.width = read.width;
    made.height = read.height;
    made.color_space = read.color_space;
    made.num_components = static_cast<int32_t>(read.components.size());
    made.max_h_samp_factor = read.max_h;
    made.max_v_samp_factor = read.max_v;
    made.progressive = read.progressive ? 1 : 0;
    made.arithmetic = read.arithmetic ? 1 : 0;
    made.exif_orientation = read.exif_orientation;
    made.warnings = read.warnings;
    bool allocated = true;
    for (size_t c = 0; c < read.components.size(); ++c) {
      const unround::JpegComponent& from = read.components[c];
      unround_jpegio_component& to = made.components[c];
      to.id = from.id;
      to.h_samp_factor = from.h;
      to.v_samp_factor = from.v;
      to.quant_table_slot = from.slot;
      to.width = from.width;
      to.height = from.height;
      to.width_in_blocks = from.blocks_across;
      to.height_in_blocks = from.blocks_down;
      std::memcpy(to.quant_table, from.table.data(), sizeof(to.quant_table));
      to.coefficients = released(from.coefficients.data(), std::max<size_t>(1, from.coefficients.size()));
      allocated = allocated && to.coefficients != nullptr;
    }
    made.icc_profile = released(read.icc_profile.data(), read.icc_profile.size());
    made.icc_profile_size = made.icc_profile != nullptr ? read.icc_profile.size() : 0;
    if (!allocated || (!read.icc_profile.empty() && made.icc_profile == nullptr)) {
      unround_jpegio_image_free(&made);
      copy_message(message, message_size, "out of memory");
      return UNROUND_JPEGIO_ERROR_MEMORY;
    }
    *image = made;
    copy_message(message, message_size, read.warnings > 0 ? read.warning : "");
    return UNROUND_JPEGIO_OK;
  } catch (...) {
    unround_jpegio_image_free(image);
    copy_message(message, message_size, "out of memory");
    return UNROUND_JPEGIO_ERROR_MEMORY;
  }
}

void unround_jpegio_image_free(unround_jpegio_image* image) {
  if (image == nullptr) return;
  for (unround_jpegio_component& c : image->components) std::free(c.coefficients);
  std::free(image->icc_profile);
  *image = unround_jpegio_image{};
}

unround_jpegio_status unround_jpegio_decode_planes(const uint8_t* data, size_t size,
                                                   const unround_jpegio_options* options, unround_jpegio_planes* planes,
                                                   char* message, size_t message_size) {
  if (planes == nullptr || (message == nullptr && message_size > 0)) return UNROUND_JPEGIO_ERROR_ARGUMENT;
  *planes = unround_jpegio_planes{};
  if (data == nullptr && size > 0) {
    copy_message(message, message_size, "the data is not addressable as one buffer");
    return UNROUND_JPEGIO_ERROR_ARGUMENT;
  }
  try {
    unround::JpegImage read;
    std::string text;
    int32_t status = UNROUND_JPEGIO_OK;
    if (!unround::read_jpeg(data, size, options != nullptr ? options->max_pixels : 0,
                            options != nullptr ? options->max_scans : 0,
                            options != nullptr && options->warnings_are_errors != 0, read, text, status)) {
      copy_message(message, message_size, text);
      return status;
    }
    std::vector<std::vector<uint8_t>> made;
    unround::jpeg_planes(read, made);
    planes->num_components = static_cast<int32_t>(made.size());
    for (size_t c = 0; c < made.size(); ++c) {
      unround_jpegio_plane& plane = planes->planes[c];
      plane.width = read.components[c].blocks_across * 8;
      plane.height = read.components[c].blocks_down * 8;
      plane.stride = plane.width;
      plane.samples = released(made[c].data(), std::max<size_t>(1, made[c].size()));
      if (plane.samples == nullptr) {
        unround_jpegio_planes_free(planes);
        copy_message(message, message_size, "out of memory");
        return UNROUND_JPEGIO_ERROR_MEMORY;
      }
    }
    copy_message(message, message_size, read.warnings > 0 ? read.warning : "");
    return UNROUND_JPEGIO_OK;
  } catch (...) {
    unround_jpegio_planes_free(planes);
    copy_message(message, message_size, "out of memory");
    return UNROUND_JPEGIO_ERROR_MEMORY;
  }
}

void unround_jpegio_planes_free(unround_jpegio_planes* planes) {
  if (planes == nullptr) return;
  for (unround_jpegio_plane& plane : planes->planes) std::free(plane.samples);
  *planes = unround_jpegio_planes{};
}

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

int32_t unround_jpegio_check_icc(const uint8_t* profile, uint64_t size, int32_t channels, char* reason,
                                 size_t reason_size) {
  try {
    std::string why;
    const bool taken = unround::check_icc(profile, size, channels, why);
    copy_message(reason, reason_size, why);
    return taken ? 1 : 0;
  } catch (...) {
    copy_message(reason, reason_size, "out of memory");
    return 0;
  }
}

unround_jpegio_status unround_jpegio_write_png(const unround_jpegio_png* png, unround_jpegio_bytes* file,
                                               char* message, size_t message_size) {
  if (file == nullptr || (message == nullptr && message_size > 0)) return UNROUND_JPEGIO_ERROR_ARGUMENT;
  *file = unround_jpegio_bytes{};
  if (png == nullptr) {
    copy_message(message, message_size, "no picture is given");
    return UNROUND_JPEGIO_ERROR_ARGUMENT;
  }
  if (png->reserved != 0) {
    copy_message(message, message_size, "the reserved field is 0");
    return UNROUND_JPEGIO_ERROR_ARGUMENT;
  }
  try {
    std::vector<uint8_t> bytes;
    std::string warning, text;
    int32_t status = UNROUND_JPEGIO_OK;
    if (!unround::png_bytes(png->samples, png->width, png->height, png->channels, png->bits, png->compression,
                            png->icc_profile, png->icc_profile_size, bytes, warning, text, status)) {
      copy_message(message, message_size, text);
      return status;
    }
    file->data = released(bytes.data(), bytes.size());
    if (file->data == nullptr) {
      copy_message(message, message_size, "out of memory");
      return UNROUND_JPEGIO_ERROR_MEMORY;
    }
    file->size = bytes.size();
    copy_message(message, message_size, warning);
    return UNROUND_JPEGIO_OK;
  } catch (...) {
    *file = unround_jpegio_bytes{};
    copy_message(message, message_size, "out of memory");
    return UNROUND_JPEGIO_ERROR_MEMORY;
  }
}

void unround_jpegio_bytes_free(unround_jpegio_bytes* bytes) {
  if (bytes == nullptr) return;
  std::free(bytes->data);
  *bytes = unround_jpegio_bytes{};
}

// The file layer under the names the reference's C interface passes it on by
// (rust/capi/include/unround.h:292-304).
int32_t unround_jpeg_abi_version(void) { return unround_jpegio_abi_version(); }

const char* unround_jpeg_libjpeg_version(void) { return unround_jpegio_libjpeg_version(); }

unround_jpegio_status unround_jpeg_read(const uint8_t* data, size_t size, const unround_jpegio_options* options,
                                        unround_jpegio_image* image, char* message, size_t message_size) {
  return unround_jpegio_read(data, size, options, image, message, message_size);
}

void unround_jpeg_image_free(unround_jpegio_image* image) { unround_jpegio_image_free(image); }

unround_jpegio_status unround_jpeg_decode_planes(const uint8_t* data, size_t size,
                                                 const unround_jpegio_options* options, unround_jpegio_planes* planes,
                                                 char* message, size_t message_size) {
  return unround_jpegio_decode_planes(data, size, options, planes, message, message_size);
}

void unround_jpeg_planes_free(unround_jpegio_planes* planes) { unround_jpegio_planes_free(planes); }

const char* unround_jpeg_libpng_version(void) { return unround_jpegio_libpng_version(); }

int32_t unround_jpeg_check_icc(const uint8_t* profile, uint64_t size, int32_t channels, char* reason,
                               size_t reason_size) {
  return unround_jpegio_check_icc(profile, size, channels, reason, reason_size);
}

}  // extern "C"
