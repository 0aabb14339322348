/* The file layer of JPEG-Unround, MI355X edition: a drop-in for the reference's C layer
 * between the files and the reconstruction (c/jpegio/include/unround/jpegio.h, ABI version
 * 3), which the reference builds on libjpeg-turbo and libpng. Here it is written on zlib
 * alone: the PNG reader and writer and the check of ICC profiles are this library's own.
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

/* The versions of the PNG layer and of the zlib it is built with
 * (c/jpegio/include/unround/jpegio.h:184-186). */
const char* unround_jpegio_libpng_version(void);

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
#endif

#endif
