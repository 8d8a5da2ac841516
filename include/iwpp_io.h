/*
 * iwpp_io.h - PGM image files for the command-line tool and the tests
 * (library: iwpp-mp_amd/libiwpp_synth.so, host code only).
 *
 * Counterpart of the reference's loadPGMToDevice (IWPP MP DQ/src/main.cu:28-125,
 * IWPP MP UQ/src/main.cu), which reads binary PGM without OpenCV.  Differences,
 * on purpose:
 *  - the size line is read as "width height", as the Netpbm format defines it
 *    (the reference reads "height width", DQ/src/main.cu:60; the two agree on
 *    the square images the reference supports);
 *  - comments ('#' to end of line) are skipped anywhere in the header;
 *  - 16-bit images (maxval 256..65535, big-endian samples) are accepted (the
 *    reference rejects maxval != 255, DQ/src/main.cu:70-73);
 *  - ASCII PGM (P2) is read as well; files are written as binary PGM (P5).
 *  - the file is read into host memory; uploading is the caller's business.
 */
#ifndef IWPP_IO_H
#define IWPP_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IWPP_IO_OK = 0,
    IWPP_IO_ERR_OPEN = -1,   /* cannot open / create the file */
    IWPP_IO_ERR_FORMAT = -2, /* not a PGM, bad header, short data */
    IWPP_IO_ERR_NOMEM = -3,  /* allocation failed */
    IWPP_IO_ERR_ARG = -4     /* bad argument */
};

/* Reads a PGM file.  On success *data points to height rows of width pixels,
 * densely packed, 1 byte per pixel when *maxval <= 255, else 2 bytes per pixel
 * in host byte order; release it with iwpp_pgm_free. */
int iwpp_pgm_read(const char* path, void** data, int* width, int* height, int* maxval);

/* Writes a binary PGM (P5).  data: height rows of `pitch` bytes holding width
 * pixels of 1 byte (maxval <= 255) or 2 bytes in host order (maxval > 255;
 * then data and pitch must be even). */
int iwpp_pgm_write(const char* path, const void* data, size_t pitch, int width, int height,
                   int maxval);

void iwpp_pgm_free(void* data);

#ifdef __cplusplus
}
#endif
#endif
