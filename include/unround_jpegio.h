/* The file layer of JPEG-Unround, MI355X edition: a drop-in for the reference's C layer
 * between the files and the reconstruction (c/jpegio/include/unround/jpegio.h, ABI version
 * 3), which the reference builds on libjpeg-turbo and libpng. Here it is written on zlib
 * alone: the JPEG reader, the PNG reader and writer and the check of ICC profiles are this
 * library's own.
 * The enum values, structure layouts and function signatures are the reference's, so that
 * its callers (the ctypes module python/src/unround/jpegio.py, the Rust crate jpegio-sys)
 * bind this library unchanged; each declaration cites the one it replaces.
 *
 * Every function returns a status and copies the message of a failure, or the first
 * warning of a success, into the caller's buffer (NUL-terminated, cut to its size; a null
 * buffer with size 0 is allowed). Nothing here keeps global state.
 */
#ifndef UNROUND_JPEGIO_H
#define UNROUND_JPEGIO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* c/jpegio/include/unround/jpegio.h:45-47 */
#define UNROUND_JPEGIO_ABI_VERSION 3
#define UNROUND_JPEGIO_MAX_COMPONENTS 4
#define UNROUND_JPEGIO_BLOCK_SIZE 64

/* c/jpegio/include/unround/jpegio.h:50-59 */
typedef int32_t unround_jpegio_status;
enum {
  UNROUND_JPEGIO_OK = 0,
  UNROUND_JPEGIO_ERROR_ARGUMENT = 1,
  UNROUND_JPEGIO_ERROR_DECODE = 2,
  UNROUND_JPEGIO_ERROR_UNSUPPORTED = 3,
  UNROUND_JPEGIO_ERROR_LIMIT = 4,
  UNROUND_JPEGIO_ERROR_MEMORY = 5,
  UNROUND_JPEGIO_ERROR_ENCODE = 6,
  UNROUND_JPEGIO_ERROR_PNG = 7,
};

/* c/jpegio/include/unround/jpegio.h:62-67 */
typedef int32_t unround_jpegio_color_space;
enum { UNROUND_JPEGIO_GRAYSCALE = 1, UNROUND_JPEGIO_YCBCR = 2, UNROUND_JPEGIO_RGB = 3 };

/* Limits of a read (c/jpegio/include/unround/jpegio.h:69-78); a null pointer for the
 * defaults. 16 bytes. */
typedef struct unround_jpegio_options {
  uint64_t max_pixels;          /* width * height; 0: 1 << 28 */
  int32_t max_scans;            /* SOS markers; 0: 500 */
  int32_t warnings_are_errors;  /* nonzero: a corrupt-data warning fails the read */
} unround_jpegio_options;

/* One component of a JPEG file (c/jpegio/include/unround/jpegio.h:80-100). 168 bytes. */
typedef struct unround_jpegio_component {
  int32_t id;
  int32_t h_samp_factor;
  int32_t v_samp_factor;
  int32_t quant_table_slot;
  uint32_t width; /* samples of the component that cover the image */
  uint32_t height;
  uint32_t width_in_blocks; /* ceil(width / 8) */
  uint32_t height_in_blocks;
  /* The table the component was quantized with, natural order: its slot's contents when
   * the component's first scan began. */
  uint16_t quant_table[UNROUND_JPEGIO_BLOCK_SIZE];
  /* height_in_blocks * width_in_blocks blocks of 64 coefficients, natural order; no dummy
   * blocks. */
  int16_t* coefficients;
} unround_jpegio_component;

/* c/jpegio/include/unround/jpegio.h:102-116. 728 bytes. */
typedef struct unround_jpegio_image {
  uint32_t width;
  uint32_t height;
  int32_t color_space;
  int32_t num_components; /* 1 or 3 */
  int32_t max_h_samp_factor;
  int32_t max_v_samp_factor;
  int32_t progressive;
  int32_t arithmetic;
  int32_t exif_orientation; /* 1 to 8, or 0 */
  int32_t warnings;         /* corrupt-data warnings while reading */
  unround_jpegio_component components[UNROUND_JPEGIO_MAX_COMPONENTS];
  uint8_t* icc_profile;
  uint64_t icc_profile_size;
} unround_jpegio_image;

/* A component decoded to 8-bit samples (c/jpegio/include/unround/jpegio.h:118-132): the
 * accurate integer inverse DCT of its coefficients, level-shifted and clamped to 0-255,
 * no smoothing, upsampling or color conversion. 24 bytes. */
typedef struct unround_jpegio_plane {
  uint32_t width;  /* width_in_blocks * 8 */
  uint32_t height; /* height_in_blocks * 8 */
  uint32_t stride;
  uint32_t reserved;
  uint8_t* samples;
} unround_jpegio_plane;

/* 104 bytes. */
typedef struct unround_jpegio_planes {
  int32_t num_components;
  int32_t reserved;
  unround_jpegio_plane planes[UNROUND_JPEGIO_MAX_COMPONENTS];
} unround_jpegio_planes;

/* A PNG file's picture as it stores it (c/jpegio/include/unround/jpegio.h:159-169): rows
 * from the top, each from the left, the samples of a pixel together, each the level the
 * file holds, 0 to 2^bits - 1. 40 bytes. */
typedef struct unround_jpegio_png_image {
  uint32_t width;
  uint32_t height;
  int32_t channels;      /* 1 grey, 2 grey and alpha, 3 RGB, 4 RGB and alpha */
  int32_t bits;          /* 1, 2, 4, 8 or 16 */
  uint16_t* samples;     /* height * width * channels levels */
  uint8_t* icc_profile;  /* the iCCP chunk's profile, or null */
  uint64_t icc_profile_size;
} unround_jpegio_png_image;

/* A picture to write as a PNG file (c/jpegio/include/unround/jpegio.h:134-157): rows from
 * the top, the samples of a pixel together (grey or RGB, then alpha, not premultiplied).
 * 48 bytes. */
typedef struct unround_jpegio_png {
  uint32_t width;
  uint32_t height;
  int32_t channels;    /* 1 to 4 */
  int32_t bits;        /* 8 or 16 */
  int32_t compression; /* zlib's level 0 to 9; -1 zlib's default */
  int32_t reserved;    /* 0 */
  const void* samples; /* uint8_t, or uint16_t in this machine's byte order */
  const uint8_t* icc_profile;
  uint64_t icc_profile_size;
} unround_jpegio_png;

/* Bytes that this layer allocated (c/jpegio/include/unround/jpegio.h:171-175). */
typedef struct unround_jpegio_bytes {
  uint8_t* data;
  uint64_t size;
} unround_jpegio_bytes;

/* c/jpegio/include/unround/jpegio.h:180-181 */
int32_t unround_jpegio_abi_version(void);

/* The JPEG reader this layer is built with (c/jpegio/include/unround/jpegio.h:180-182): this
 * library's own, from ITU-T T.81. */
const char* unround_jpegio_libjpeg_version(void);

/* The versions of the PNG layer and of the zlib it is built with
 * (c/jpegio/include/unround/jpegio.h:184-186). */
const char* unround_jpegio_libpng_version(void);

/* Reads the coefficients and metadata of a JPEG file in memory
 * (c/jpegio/include/unround/jpegio.h:188-199): baseline, extended and progressive Huffman
 * files of 8 bits, grey or three components, with restart intervals; arithmetic coding,
 * 12-bit, lossless, hierarchical and CMYK files are UNROUND_JPEGIO_ERROR_UNSUPPORTED. On
 * success image owns what it points to until unround_jpegio_image_free(), and message holds
 * the first corrupt-data warning or an empty string. */
unround_jpegio_status unround_jpegio_read(const uint8_t* data, size_t size, const unround_jpegio_options* options,
                                          unround_jpegio_image* image, char* message, size_t message_size);

/* c/jpegio/include/unround/jpegio.h:201-203 */
void unround_jpegio_image_free(unround_jpegio_image* image);

/* The component planes of a JPEG file in memory, with unround_jpegio_read's contract
 * (c/jpegio/include/unround/jpegio.h:205-209). */
unround_jpegio_status unround_jpegio_decode_planes(const uint8_t* data, size_t size,
                                                   const unround_jpegio_options* options, unround_jpegio_planes* planes,
                                                   char* message, size_t message_size);

/* c/jpegio/include/unround/jpegio.h:211-213 */
void unround_jpegio_planes_free(unround_jpegio_planes* planes);

/* Reads a PNG file in memory as the levels it stores (c/jpegio/include/unround/jpegio.h:
 * 239-250): grey, grey and alpha, RGB or RGB and alpha, of every depth PNG has,
 * interlaced or not, with the profile of its iCCP chunk. A palette file is refused
 * (UNROUND_JPEGIO_ERROR_UNSUPPORTED); more than max_pixels pixels (0: 1 << 28) is
 * UNROUND_JPEGIO_ERROR_LIMIT; a file this layer cannot read is UNROUND_JPEGIO_ERROR_PNG.
 * On success image owns what it points to until unround_jpegio_png_image_free(), and
 * message holds the first warning (a tRNS chunk's transparent colour, not taken) or an
 * empty string; on failure image is left empty and message says why. */
unround_jpegio_status unround_jpegio_read_png(const uint8_t* data, size_t size, uint64_t max_pixels,
                                              unround_jpegio_png_image* image, char* message, size_t message_size);

/* c/jpegio/include/unround/jpegio.h:252-254 */
void unround_jpegio_png_image_free(unround_jpegio_png_image* image);

/* Whether an ICC profile goes with a picture of the channels (1, 2: grey; 3, 4: RGB): at
 * least 132 bytes, the length its header gives, a multiple of 4 from version 4 on, the
 * signature 'acsp', its tags within it, a rendering intent below 0xFFFF, a class neither
 * abstract nor a device link, the picture's data colour space and XYZ or Lab as its
 * connection space (c/jpegio/include/unround/jpegio.h:212-223). 1 where it does; 0 with
 * the reason where it does not. */
int32_t unround_jpegio_check_icc(const uint8_t* profile, uint64_t size, int32_t channels, char* reason,
                                 size_t reason_size);

/* Writes a picture as a PNG file into memory (c/jpegio/include/unround/jpegio.h:225-233):
 * 8 or 16 bits, grey or RGB with or without alpha, not interlaced, each row by the filter
 * of the least sum, with the ICC profile where unround_jpegio_check_icc takes it (else the
 * message says why). file owns the bytes until unround_jpegio_bytes_free(). */
unround_jpegio_status unround_jpegio_write_png(const unround_jpegio_png* png, unround_jpegio_bytes* file,
                                               char* message, size_t message_size);

/* c/jpegio/include/unround/jpegio.h:235-237 */
void unround_jpegio_bytes_free(unround_jpegio_bytes* bytes);

#ifdef __cplusplus
}
static_assert(sizeof(unround_jpegio_options) == 16, "the layout of unround_jpegio_options");
static_assert(sizeof(unround_jpegio_component) == 168, "the layout of unround_jpegio_component");
static_assert(sizeof(unround_jpegio_image) == 728, "the layout of unround_jpegio_image");
static_assert(sizeof(unround_jpegio_plane) == 24, "the layout of unround_jpegio_plane");
static_assert(sizeof(unround_jpegio_planes) == 104, "the layout of unround_jpegio_planes");
static_assert(sizeof(unround_jpegio_png_image) == 40, "the layout of unround_jpegio_png_image");
static_assert(sizeof(unround_jpegio_png) == 48, "the layout of unround_jpegio_png");
#endif

#endif
