/* The C interface of JPEG-Unround, MI355X edition.
 *
 * A drop-in for the reference's C ABI (rust/capi/include/unround.h, ABI version 5) for
 * its reconstruction path: the TV/TGV primal-dual solve of components given as arrays,
 * computed by HIP kernels on an AMD Instinct GPU (gfx950). The declarations, enum values
 * and structure layouts are the reference's, so that its callers (the ctypes module
 * python/src/unround/native.py, a cgo or JNI stub) bind this library unchanged; each
 * declaration cites the one it replaces.
 *
 * Settings are made from options under the command line's names ("--mu", "0.01",
 * "--method=tv"). Every function that computes returns a status, and writes what went
 * wrong into a message buffer that the caller gives (NUL-terminated, cut at a UTF-8
 * boundary; a null buffer is allowed). No C++ exception crosses this interface.
 * Different settings and results can be used on different threads at once, and one can
 * be read by several.
 *
 * unround_decode reads JPEG files (this library's own reader of baseline, extended and
 * progressive Huffman files) and reconstructs their coefficients on the GPU, and reads PNG
 * files (the file layer, include/unround_jpegio.h, on zlib) and reconstructs them within
 * the intervals of their rounded samples. unround_project projects a picture onto a JPEG
 * file's intervals on the GPU. Arithmetic-coded, 12-bit, lossless, hierarchical and CMYK
 * JPEG files are UNROUND_ERROR_UNSUPPORTED, and so is unround_main (the command line,
 * which returns 2). The writers of TIFF, PNG and PNM, with the rounding and blue-noise
 * dither of docs/math.md 10, and unround_orient are this library's own.
 */
#ifndef UNROUND_H
#define UNROUND_H

#include <stddef.h>
#include <stdint.h>

#include "unround_jpegio.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rust/capi/include/unround.h:41 */
#define UNROUND_ABI_VERSION 5

/* The color spaces of an input (c/jpegio/include/unround/jpegio.h:63-67). */
#define UNROUND_COLOR_GRAYSCALE 1
#define UNROUND_COLOR_YCBCR 2
#define UNROUND_COLOR_RGB 3

/* rust/capi/include/unround.h:44-52 */
typedef int32_t unround_status;
enum {
  UNROUND_OK = 0,
  UNROUND_ERROR_ARGUMENT = 1,
  UNROUND_ERROR_READ = 2,
  UNROUND_ERROR_UNSUPPORTED = 3,
  UNROUND_ERROR_OPTIONS = 4,
  UNROUND_ERROR_IO = 5,
  UNROUND_ERROR_INTERNAL = 6,
};

/* rust/capi/include/unround.h:54-57 */
int32_t unround_abi_version(void);
const char* unround_version(void);

/* ---------------------------------------------------------------------------------
 * Settings (rust/capi/include/unround.h:59-76) */

typedef struct unround_settings unround_settings;

/* Settings from count options; null where they are refused, with every refusal in the
 * message. Besides the reference's options, --precision single|double|mixed chooses the
 * arithmetic of the iterations (binary32, binary64, or binary32 with the dual fields
 * stored in bfloat16); single by default. */
unround_settings* unround_settings_new(size_t count, const char* const* options, char* message,
                                       size_t message_size);
void unround_settings_free(unround_settings* settings);
/* The options that give the settings, "--name=value" separated by spaces; returns the
 * length of the whole text without its NUL, 0 for null settings. */
size_t unround_settings_options(const unround_settings* settings, char* buffer, size_t size);

/* ---------------------------------------------------------------------------------
 * Reconstruction (rust/capi/include/unround.h:78-140) */

/* rust/capi/include/unround.h:81-91: 64 bytes. */
typedef struct unround_record {
  uint64_t iteration;
  double gap;
  double primal;
  double dual;
  uint64_t channels;
  uint64_t height;
  uint64_t width;
  const double* canvas;
} unround_record;

/* Called with each record after the first; a nonzero return stops the solver. */
typedef int32_t (*unround_observer)(void* user, const unround_record* record);

/* rust/capi/include/unround.h:99-108: 152 bytes, quant_table at offset 24. */
typedef struct unround_component {
  const int16_t* levels;
  uint32_t rows;
  uint32_t columns;
  int32_t h_samp_factor;
  int32_t v_samp_factor;
  uint16_t quant_table[64];
} unround_component;

/* rust/capi/include/unround.h:110-118: 24 bytes, components at offset 16. */
typedef struct unround_input {
  uint32_t height;
  uint32_t width;
  int32_t color_space;
  uint32_t count;
  const unround_component* components;
} unround_input;

typedef struct unround_result unround_result;

/* rust/capi/include/unround.h:122-126 */
typedef int32_t unround_format;
enum { UNROUND_FORMAT_JPEG = 1, UNROUND_FORMAT_PNG = 2 };

/* rust/capi/include/unround.h:128-134: a JPEG file (known by its signature) read to its
 * levels, tables and sampling and solved as unround_solve solves them; or a PNG file
 * within the intervals of its rounded samples, by TV or TGV with --mu and --slack alone of
 * the data term, its colours together and its alpha alone. A JPEG coding this library does
 * not read (arithmetic, 12-bit, lossless, hierarchical, CMYK) is
 * UNROUND_ERROR_UNSUPPORTED. */
unround_status unround_decode(const uint8_t* data, size_t size, const unround_settings* settings,
                              unround_observer observer, void* user, unround_result** result, char* message,
                              size_t message_size);

/* rust/capi/include/unround.h:136-140; rust/capi/src/lib.rs:669-694. Reconstructs
 * components given as arrays on the GPU. */
unround_status unround_solve(const unround_input* input, const unround_settings* settings, unround_observer observer,
                             void* user, unround_result** result, char* message, size_t message_size);

/* rust/capi/include/unround.h:142-150: a picture (height x width x channels
 * interleaved, RGB for a YCbCr file) projected onto the intervals of a JPEG file's
 * levels on the GPU, each coefficient's distance outside its interval kept as the
 * result's excess (docs/math.md 12). */
unround_status unround_project(const uint8_t* data, size_t size, const unround_settings* settings,
                               const double* picture, uint64_t height, uint64_t width, uint64_t channels,
                               unround_result** result, char* message, size_t message_size);

/* ---------------------------------------------------------------------------------
 * Results (rust/capi/include/unround.h:152-244) */

void unround_result_free(unround_result* result);
const double* unround_result_picture(const unround_result* result, uint64_t* height, uint64_t* width,
                                     uint64_t* channels);
const double* unround_result_planes(const unround_result* result, uint64_t* count);
int32_t unround_result_color_space(const unround_result* result);
const double* unround_result_canvas(const unround_result* result, uint64_t* channels, uint64_t* height,
                                    uint64_t* width);
uint64_t unround_result_components(const unround_result* result);
const double* unround_result_coefficients(const unround_result* result, uint64_t component, uint64_t* count);
const double* unround_result_problem(const unround_result* result, uint64_t component, int32_t which, uint64_t* count);
unround_status unround_result_component(const unround_result* result, uint64_t component, uint64_t* rows,
                                        uint64_t* columns, uint64_t* down, uint64_t* across);
int32_t unround_result_solver(const unround_result* result, uint64_t* iterations, int32_t* stop);
const uint64_t* unround_result_record_iterations(const unround_result* result, uint64_t* count);
const double* unround_result_record_values(const unround_result* result, int32_t which, uint64_t* count);
int32_t unround_result_alpha_solver(const unround_result* result, uint64_t* iterations, int32_t* stop);
const uint64_t* unround_result_alpha_record_iterations(const unround_result* result, uint64_t* count);
const double* unround_result_alpha_record_values(const unround_result* result, int32_t which, uint64_t* count);
const uint8_t* unround_result_icc_profile(const unround_result* result, uint64_t* size);
int32_t unround_result_exif_orientation(const unround_result* result);
int32_t unround_result_format(const unround_result* result);
int32_t unround_result_bits(const unround_result* result);
const char* unround_result_warning(const unround_result* result);
const double* unround_result_excess(const unround_result* result, uint64_t component, uint64_t* count);

/* ---------------------------------------------------------------------------------
 * Output and conversion (rust/capi/include/unround.h:235-290) */

/* rust/capi/include/unround.h:246-254 */
typedef int32_t unround_dither;
enum { UNROUND_DITHER_NONE = 0, UNROUND_DITHER_BLUE_NOISE = 1, UNROUND_DITHER_BLUE_NOISE_TOGETHER = 2 };

unround_status unround_tiff(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t ycbcr,
                            const uint8_t* icc, uint64_t icc_size, uint8_t** bytes, uint64_t* size, char* message,
                            size_t message_size);
unround_status unround_png(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t sixteen,
                           int32_t dither, int32_t compression, const uint8_t* icc, uint64_t icc_size, uint8_t** bytes,
                           uint64_t* size, char* message, size_t message_size);
unround_status unround_pnm(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t sixteen,
                           int32_t dither, uint8_t** bytes, uint64_t* size, char* message, size_t message_size);
void unround_bytes_free(uint8_t* bytes, uint64_t size);
unround_status unround_orient(const double* samples, uint64_t height, uint64_t width, uint64_t channels,
                              int32_t orientation, double* turned, uint64_t* turned_height, uint64_t* turned_width,
                              char* message, size_t message_size);
/* JFIF's RGB of Y, Cb and Cr planes, and back (docs/math.md:995-1040). */
unround_status unround_to_rgb(const double* planes, uint64_t height, uint64_t width, double* picture);
unround_status unround_to_ycbcr(const double* picture, uint64_t height, uint64_t width, double* planes);
int32_t unround_main(size_t count, const char* const* arguments);

/* ---------------------------------------------------------------------------------
 * The file layer through this library (rust/capi/include/unround.h:292-304): the
 * functions of include/unround_jpegio.h under the names the reference passes them on by. */
int32_t unround_jpeg_abi_version(void);
const char* unround_jpeg_libjpeg_version(void);
const char* unround_jpeg_libpng_version(void);
int32_t unround_jpeg_check_icc(const uint8_t* profile, uint64_t size, int32_t channels, char* reason,
                               size_t reason_size);
unround_jpegio_status unround_jpeg_read(const uint8_t* data, size_t size, const unround_jpegio_options* options,
                                        unround_jpegio_image* image, char* message, size_t message_size);
void unround_jpeg_image_free(unround_jpegio_image* image);
unround_jpegio_status unround_jpeg_decode_planes(const uint8_t* data, size_t size,
                                                 const unround_jpegio_options* options, unround_jpegio_planes* planes,
                                                 char* message, size_t message_size);
void unround_jpeg_planes_free(unround_jpegio_planes* planes);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(unround_record) == 64, "the layout of unround_record");
static_assert(sizeof(unround_component) == 152, "the layout of unround_component");
static_assert(offsetof(unround_component, quant_table) == 24, "the layout of unround_component");
static_assert(sizeof(unround_input) == 24, "the layout of unround_input");
static_assert(offsetof(unround_input, components) == 16, "the layout of unround_input");
#else
_Static_assert(sizeof(unround_record) == 64, "the layout of unround_record");
_Static_assert(sizeof(unround_component) == 152, "the layout of unround_component");
_Static_assert(offsetof(unround_component, quant_table) == 24, "the layout of unround_component");
_Static_assert(sizeof(unround_input) == 24, "the layout of unround_input");
_Static_assert(offsetof(unround_input, components) == 16, "the layout of unround_input");
#endif

#endif
