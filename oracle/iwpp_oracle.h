/*
 * CPU oracle for IWPP grayscale morphological reconstruction by dilation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product path (libiwpp.so) never links or calls it.
 *
 * The reference (oliveira-mathias/IWPP-MP) ships no CPU reconstruction; its
 * GPU kernels compute the fixed point of the update rule
 *   "if J(q) > J(p) and J(p) < I(p) then J(p) = min(J(q), I(p))"
 * (IWPP MP/src/MorphologicalReconstruction.cu:42-47) over 4-neighbours, with
 * pixels outside the image absent (zero halos, same file :252-312), after
 * clipping marker to mask (same file :827-836).  Because reconstruction by
 * dilation is the unique fixed point of that rule, the oracle computes it
 * with Vincent's sequential hybrid (raster, anti-raster, FIFO; L. Vincent,
 * IEEE TIP 2(2), 1993), for 4- and 8-connectivity, uint8 and uint16, any
 * width x height, any row pitch.
 *
 * Images are passed as byte pointers plus row pitches in bytes.
 * dtype: 1 = uint8, 2 = uint16.  conn: 4 or 8.
 * All functions return 0 on success and a negative value on bad arguments.
 */
#ifndef IWPP_ORACLE_H
#define IWPP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* J = min(J, I) per pixel (IWPP MP/src/MorphologicalReconstruction.cu:827-836). */
int oracle_clip(void* marker, const void* mask, size_t pitch_j, size_t pitch_i,
                int width, int height, int dtype);

/* Full reconstruction in place: clip, raster, anti-raster, FIFO propagation. */
int oracle_reconstruct(void* marker, const void* mask, size_t pitch_j, size_t pitch_i,
                       int width, int height, int dtype, int conn);

/* One row pass, forward (left to right) then backward, over every row:
 *   out[x] = max(J[x], min(out[x-1], I[x]))   (forward; backward mirrors it).
 * The 1-D sequential form of the reference's X sweeps
 * (IWPP MP DQ/src/MorphologicalReconstruction.cu:817-1040). */
int oracle_pass_x(void* marker, const void* mask, size_t pitch_j, size_t pitch_i,
                  int width, int height, int dtype);

/* One column pass, blocked and out of place.  The image is cut into blocks of
 * block_w columns x block_h rows.  Inside a block, rows are swept in order
 * (down: top to bottom; up: bottom to top):
 *   4-conn: out[y][x] = max(src[y][x], min(prev[x], I[y][x]))
 *   8-conn: out[y][x] = max(src[y][x], min(max(prev[x-1], prev[x], prev[x+1]), I[y][x]))
 * where prev is the block's previous output row; neighbours that lie outside
 * the block (the halo row above/below, the columns left/right) are read from
 * src; neighbours outside the image are absent.  dir: +1 down, -1 up.
 * block_w >= width and block_h >= height give the classic full-length pass. */
int oracle_pass_y(const void* src, void* dst, const void* mask, size_t pitch_s,
                  size_t pitch_d, size_t pitch_i, int width, int height, int dtype,
                  int conn, int dir, int block_w, int block_h);

/* One raster pass, blocked and out of place: the blocks of oracle_pass_y, and
 * after each row's column step the row is scanned across its block, forward
 *   out[y][x] = max(out[y][x], min(out[y][x-1], I[y][x]))
 * from x = bx0 (out[y][bx0-1] read from src, 0 outside the image), then
 * backward the same way from x = bx1-1 (out[y][bx1] from src). */
int oracle_pass_raster(const void* src, void* dst, const void* mask, size_t pitch_s,
                       size_t pitch_d, size_t pitch_i, int width, int height, int dtype,
                       int conn, int dir, int block_w, int block_h);

/* Stability test per megapixel block of mp_w x mp_h pixels.  flags[b] = 1 when
 * some pixel p of block b has a neighbour q with min(J(q), I(p)) > J(p).
 * flags has ceil(h/mp_h) * ceil(w/mp_w) entries, row-major.  Returns the
 * number of unstable blocks. */
long oracle_unstable_blocks(const void* marker, const void* mask, size_t pitch_j,
                            size_t pitch_i, int width, int height, int dtype, int conn,
                            int mp_w, int mp_h, uint8_t* flags);

#ifdef __cplusplus
}
#endif
#endif
