/*
 * CPU oracle for IWPP morphological reconstruction by dilation.
 * TEST INFRASTRUCTURE ONLY: the checker used by tests, the smoke check and
 * the benchmark's CPU-baseline leg; never part of the product path.
 *
 * Restates, on the CPU, what the reference's GPU code computes:
 *   - clip:  IWPP MP/src/MorphologicalReconstruction.cu:827-836
 *   - update rule and zero (absent) halos: same file :42-47, :252-312
 *   - 1-D sweep recurrence s = min(max(prev, old), mask):
 *     IWPP MP DQ/src/MorphologicalReconstruction.cu:861, :934
 * and computes the unique fixed point with Vincent's hybrid algorithm
 * (raster pass, anti-raster pass seeding a FIFO, FIFO propagation;
 * L. Vincent, "Morphological grayscale reconstruction in image analysis",
 * IEEE Trans. Image Processing 2(2), 1993).
 */
#include "iwpp_oracle.h"

#include <stdlib.h>
#include <string.h>

#define ROWP(T, base, pitch, y) ((T*)((uint8_t*)(base) + (size_t)(y) * (pitch)))
#define CROWP(T, base, pitch, y) ((const T*)((const uint8_t*)(base) + (size_t)(y) * (pitch)))
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))

static int check_args(const void* a, const void* b, size_t pa, size_t pb, int w, int h,
                      int dtype) {
    if (!a || !b || w <= 0 || h <= 0) return -1;
    if (dtype != 1 && dtype != 2) return -2;
    if (pa < (size_t)w * dtype || pb < (size_t)w * dtype) return -3;
    return 0;
}

/* ---- growable FIFO of pixel indices (uint32: images up to 2^32-1 px) ---- */
typedef struct {
    uint32_t* buf;
    size_t cap, head, count;
} fifo_t;

static int fifo_init(fifo_t* f, size_t cap) {
    f->cap = cap < 1024 ? 1024 : cap;
    f->head = f->count = 0;
    f->buf = (uint32_t*)malloc(f->cap * sizeof(uint32_t));
    return f->buf ? 0 : -1;
}
static int fifo_push(fifo_t* f, uint32_t v) {
    if (f->count == f->cap) {
        size_t ncap = f->cap * 2;
        uint32_t* nb = (uint32_t*)malloc(ncap * sizeof(uint32_t));
        if (!nb) return -1;
        for (size_t i = 0; i < f->count; ++i) nb[i] = f->buf[(f->head + i) % f->cap];
        free(f->buf);
        f->buf = nb;
        f->cap = ncap;
        f->head = 0;
    }
    f->buf[(f->head + f->count) % f->cap] = v;
    f->count++;
    return 0;
}
static uint32_t fifo_pop(fifo_t* f) {
    uint32_t v = f->buf[f->head];
    f->head = (f->head + 1) % f->cap;
    f->count--;
    return v;
}

/* Neighbour offsets: the first half (raster-previous, N+) then the mirror (N-). */
static const int NB8_DY[8] = {0, -1, -1, -1, 0, 1, 1, 1};
static const int NB8_DX[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
static const int NB4_DY[4] = {0, -1, 0, 1};
static const int NB4_DX[4] = {-1, 0, 1, 0};

#define DEFINE_ORACLE(T, SUF)                                                                  \
    static void clip_##SUF(void* jb, const void* ib, size_t pj, size_t pi, int w, int h) {      \
        for (int y = 0; y < h; ++y) {                                                           \
            T* J = ROWP(T, jb, pj, y);                                                          \
            const T* I = CROWP(T, ib, pi, y);                                                   \
            for (int x = 0; x < w; ++x) J[x] = MIN(J[x], I[x]);                                 \
        }                                                                                       \
    }                                                                                           \
                                                                                                \
    static int vincent_##SUF(void* jb, const void* ib, size_t pj, size_t pi, int w, int h,      \
                             int conn) {                                                        \
        const int* DY = conn == 8 ? NB8_DY : NB4_DY;                                            \
        const int* DX = conn == 8 ? NB8_DX : NB4_DX;                                            \
        const int nn = conn == 8 ? 8 : 4, half = nn / 2;                                        \
        clip_##SUF(jb, ib, pj, pi, w, h);                                                       \
        /* raster pass: J(p) = min(max(J(p), J(N+(p))), I(p)) */                                \
        for (int y = 0; y < h; ++y) {                                                           \
            T* J = ROWP(T, jb, pj, y);                                                          \
            const T* I = CROWP(T, ib, pi, y);                                                   \
            const T* Ju = y > 0 ? ROWP(T, jb, pj, y - 1) : NULL;                                \
            for (int x = 0; x < w; ++x) {                                                       \
                T m = J[x];                                                                     \
                if (x > 0) m = MAX(m, J[x - 1]);                                                \
                if (Ju) {                                                                       \
                    m = MAX(m, Ju[x]);                                                          \
                    if (conn == 8) {                                                            \
                        if (x > 0) m = MAX(m, Ju[x - 1]);                                       \
                        if (x + 1 < w) m = MAX(m, Ju[x + 1]);                                   \
                    }                                                                           \
                }                                                                               \
                J[x] = MIN(m, I[x]);                                                            \
            }                                                                                   \
        }                                                                                       \
        fifo_t f;                                                                               \
        if (fifo_init(&f, (size_t)w * 4) != 0) return -10;                                      \
        /* anti-raster pass, seeding the FIFO with pixels that can still raise an              \
           N- neighbour */                                                                      \
        for (int y = h - 1; y >= 0; --y) {                                                      \
            T* J = ROWP(T, jb, pj, y);                                                          \
            const T* I = CROWP(T, ib, pi, y);                                                   \
            T* Jd = y + 1 < h ? ROWP(T, jb, pj, y + 1) : NULL;                                  \
            const T* Id = y + 1 < h ? CROWP(T, ib, pi, y + 1) : NULL;                           \
            for (int x = w - 1; x >= 0; --x) {                                                  \
                T m = J[x];                                                                     \
                if (x + 1 < w) m = MAX(m, J[x + 1]);                                            \
                if (Jd) {                                                                       \
                    m = MAX(m, Jd[x]);                                                          \
                    if (conn == 8) {                                                            \
                        if (x + 1 < w) m = MAX(m, Jd[x + 1]);                                   \
                        if (x > 0) m = MAX(m, Jd[x - 1]);                                       \
                    }                                                                           \
                }                                                                               \
                const T v = MIN(m, I[x]);                                                       \
                J[x] = v;                                                                       \
                int push = 0;                                                                   \
                if (x + 1 < w && J[x + 1] < v && J[x + 1] < I[x + 1]) push = 1;                 \
                if (!push && Jd) {                                                              \
                    if (Jd[x] < v && Jd[x] < Id[x]) push = 1;                                   \
                    if (!push && conn == 8) {                                                   \
                        if (x + 1 < w && Jd[x + 1] < v && Jd[x + 1] < Id[x + 1]) push = 1;      \
                        if (!push && x > 0 && Jd[x - 1] < v && Jd[x - 1] < Id[x - 1]) push = 1; \
                    }                                                                           \
                }                                                                               \
                if (push && fifo_push(&f, (uint32_t)((size_t)y * w + x)) != 0) {                \
                    free(f.buf);                                                                \
                    return -10;                                                                 \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        /* FIFO propagation over the full neighbourhood */                                      \
        while (f.count) {                                                                       \
            const uint32_t p = fifo_pop(&f);                                                    \
            const int py = (int)(p / (uint32_t)w), px = (int)(p % (uint32_t)w);                 \
            const T vp = ROWP(T, jb, pj, py)[px];                                               \
            for (int k = 0; k < nn; ++k) {                                                      \
                const int qy = py + DY[k], qx = px + DX[k];                                     \
                if (qy < 0 || qy >= h || qx < 0 || qx >= w) continue;                           \
                T* Jq = ROWP(T, jb, pj, qy) + qx;                                               \
                const T iq = CROWP(T, ib, pi, qy)[qx];                                          \
                if (*Jq < vp && *Jq != iq) {                                                    \
                    *Jq = MIN(vp, iq);                                                          \
                    if (fifo_push(&f, (uint32_t)((size_t)qy * w + qx)) != 0) {                  \
                        free(f.buf);                                                            \
                        return -10;                                                             \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        (void)half;                                                                             \
        free(f.buf);                                                                            \
        return 0;                                                                               \
    }                                                                                           \
                                                                                                \
    static void pass_x_##SUF(void* jb, const void* ib, size_t pj, size_t pi, int w, int h) {    \
        for (int y = 0; y < h; ++y) {                                                           \
            T* J = ROWP(T, jb, pj, y);                                                          \
            const T* I = CROWP(T, ib, pi, y);                                                   \
            for (int x = 1; x < w; ++x) J[x] = MAX(J[x], MIN(J[x - 1], I[x]));                  \
            for (int x = w - 2; x >= 0; --x) J[x] = MAX(J[x], MIN(J[x + 1], I[x]));             \
        }                                                                                       \
    }                                                                                           \
                                                                                                \
    static void pass_y_##SUF(const void* sb, void* db, const void* ib, size_t ps, size_t pd,    \
                             size_t pi, int w, int h, int conn, int dir, int bw, int bh) {      \
        for (int by0 = 0; by0 < h; by0 += bh) {                                                 \
            const int by1 = MIN(h, by0 + bh);                                                   \
            for (int bx0 = 0; bx0 < w; bx0 += bw) {                                             \
                const int bx1 = MIN(w, bx0 + bw);                                               \
                const int ystart = dir > 0 ? by0 : by1 - 1;                                     \
                const int yend = dir > 0 ? by1 : by0 - 1;                                       \
                for (int y = ystart; y != yend; y += dir) {                                     \
                    const int yp = y - dir; /* previous row in sweep order */                   \
                    const int prev_in_block = yp >= by0 && yp < by1;                            \
                    const int prev_in_image = yp >= 0 && yp < h;                                \
                    const T* S = CROWP(T, sb, ps, y);                                           \
                    const T* I = CROWP(T, ib, pi, y);                                           \
                    T* D = ROWP(T, db, pd, y);                                                  \
                    const T* Pd = prev_in_image ? CROWP(T, db, pd, yp) : NULL;                  \
                    const T* Ps = prev_in_image ? CROWP(T, sb, ps, yp) : NULL;                  \
                    for (int x = bx0; x < bx1; ++x) {                                           \
                        T m = 0;                                                                \
                        if (prev_in_image) {                                                    \
                            m = prev_in_block ? Pd[x] : Ps[x];                                  \
                            if (conn == 8) {                                                    \
                                if (x - 1 >= 0) {                                               \
                                    T l = (prev_in_block && x - 1 >= bx0) ? Pd[x - 1]           \
                                                                          : Ps[x - 1];          \
                                    m = MAX(m, l);                                              \
                                }                                                               \
                                if (x + 1 < w) {                                                \
                                    T r = (prev_in_block && x + 1 < bx1) ? Pd[x + 1]            \
                                                                         : Ps[x + 1];           \
                                    m = MAX(m, r);                                              \
                                }                                                               \
                            }                                                                   \
                        }                                                                       \
                        D[x] = MAX(S[x], MIN(m, I[x]));                                         \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
    }                                                                                           \
                                                                                                \
    static long unstable_##SUF(const void* jb, const void* ib, size_t pj, size_t pi, int w,     \
                               int h, int conn, int mw, int mh, uint8_t* flags) {               \
        const int nbx = (w + mw - 1) / mw, nby = (h + mh - 1) / mh;                             \
        const int* DY = conn == 8 ? NB8_DY : NB4_DY;                                            \
        const int* DX = conn == 8 ? NB8_DX : NB4_DX;                                            \
        const int nn = conn == 8 ? 8 : 4;                                                       \
        memset(flags, 0, (size_t)nbx * nby);                                                    \
        long count = 0;                                                                         \
        for (int y = 0; y < h; ++y) {                                                           \
            const T* J = CROWP(T, jb, pj, y);                                                   \
            const T* I = CROWP(T, ib, pi, y);                                                   \
            for (int x = 0; x < w; ++x) {                                                       \
                uint8_t* fl = &flags[(size_t)(y / mh) * nbx + x / mw];                          \
                if (*fl || J[x] == I[x]) continue;                                              \
                for (int k = 0; k < nn; ++k) {                                                  \
                    const int qy = y + DY[k], qx = x + DX[k];                                   \
                    if (qy < 0 || qy >= h || qx < 0 || qx >= w) continue;                       \
                    const T jq = CROWP(T, jb, pj, qy)[qx];                                      \
                    if (MIN(jq, I[x]) > J[x]) {                                                 \
                        *fl = 1;                                                                \
                        ++count;                                                                \
                        break;                                                                  \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
        return count;                                                                           \
    }

DEFINE_ORACLE(uint8_t, u8)
DEFINE_ORACLE(uint16_t, u16)

/* Raster pass over blocks (bw columns x bh rows), out of place: in each block
 * the rows are taken in sweep order; a row first gets the column step of
 * pass_y (from the block's previous output row; neighbours outside the block
 * from src), then a forward and a backward scan across the block's columns,
 * entering from src's pixels just left and right of the block in the same
 * row (0 outside the image).  Vincent's raster scan (1993) with the row
 * scanned both ways, restricted to one block; the pixels it reads outside a
 * block are lower bounds of the reconstruction, so any order of such passes
 * has the same fixed point. */
#define DEFINE_RASTER(T, SUF)                                                                   \
    static void pass_raster_##SUF(const void* sb, void* db, const void* ib, size_t ps,          \
                                  size_t pd, size_t pi, int w, int h, int conn, int dir,        \
                                  int bw, int bh) {                                             \
        for (int by0 = 0; by0 < h; by0 += bh) {                                                 \
            const int by1 = MIN(h, by0 + bh);                                                   \
            for (int bx0 = 0; bx0 < w; bx0 += bw) {                                             \
                const int bx1 = MIN(w, bx0 + bw);                                               \
                const int ystart = dir > 0 ? by0 : by1 - 1;                                     \
                const int yend = dir > 0 ? by1 : by0 - 1;                                       \
                for (int y = ystart; y != yend; y += dir) {                                     \
                    const int yp = y - dir;                                                     \
                    const int prev_in_block = yp >= by0 && yp < by1;                            \
                    const int prev_in_image = yp >= 0 && yp < h;                                \
                    const T* S = CROWP(T, sb, ps, y);                                           \
                    const T* I = CROWP(T, ib, pi, y);                                           \
                    T* D = ROWP(T, db, pd, y);                                                  \
                    const T* Pd = prev_in_image ? CROWP(T, db, pd, yp) : NULL;                  \
                    const T* Ps = prev_in_image ? CROWP(T, sb, ps, yp) : NULL;                  \
                    for (int x = bx0; x < bx1; ++x) {                                           \
                        T m = 0;                                                                \
                        if (prev_in_image) {                                                    \
                            m = prev_in_block ? Pd[x] : Ps[x];                                  \
                            if (conn == 8) {                                                    \
                                if (x - 1 >= 0)                                                 \
                                    m = MAX(m, (prev_in_block && x - 1 >= bx0) ? Pd[x - 1]      \
                                                                               : Ps[x - 1]);    \
                                if (x + 1 < w)                                                  \
                                    m = MAX(m, (prev_in_block && x + 1 < bx1) ? Pd[x + 1]       \
                                                                              : Ps[x + 1]);     \
                            }                                                                   \
                        }                                                                       \
                        D[x] = MAX(S[x], MIN(m, I[x]));                                         \
                    }                                                                           \
                    T c = bx0 > 0 ? S[bx0 - 1] : 0;                                             \
                    for (int x = bx0; x < bx1; ++x) {                                           \
                        D[x] = MAX(D[x], MIN(c, I[x]));                                         \
                        c = D[x];                                                               \
                    }                                                                           \
                    c = bx1 < w ? S[bx1] : 0;                                                   \
                    for (int x = bx1 - 1; x >= bx0; --x) {                                      \
                        D[x] = MAX(D[x], MIN(c, I[x]));                                         \
                        c = D[x];                                                               \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
    }

DEFINE_RASTER(uint8_t, u8)
DEFINE_RASTER(uint16_t, u16)

int oracle_clip(void* marker, const void* mask, size_t pj, size_t pi, int w, int h, int dtype) {
    int rc = check_args(marker, mask, pj, pi, w, h, dtype);
    if (rc) return rc;
    if (dtype == 1) clip_u8(marker, mask, pj, pi, w, h);
    else clip_u16(marker, mask, pj, pi, w, h);
    return 0;
}

int oracle_reconstruct(void* marker, const void* mask, size_t pj, size_t pi, int w, int h,
                       int dtype, int conn) {
    int rc = check_args(marker, mask, pj, pi, w, h, dtype);
    if (rc) return rc;
    if (conn != 4 && conn != 8) return -4;
    if ((uint64_t)w * (uint64_t)h > 0xFFFFFFFFull) return -5;
    return dtype == 1 ? vincent_u8(marker, mask, pj, pi, w, h, conn)
                      : vincent_u16(marker, mask, pj, pi, w, h, conn);
}

int oracle_pass_x(void* marker, const void* mask, size_t pj, size_t pi, int w, int h,
                  int dtype) {
    int rc = check_args(marker, mask, pj, pi, w, h, dtype);
    if (rc) return rc;
    if (dtype == 1) pass_x_u8(marker, mask, pj, pi, w, h);
    else pass_x_u16(marker, mask, pj, pi, w, h);
    return 0;
}

int oracle_pass_y(const void* src, void* dst, const void* mask, size_t ps, size_t pd,
                  size_t pi, int w, int h, int dtype, int conn, int dir, int bw, int bh) {
    int rc = check_args(src, mask, ps, pi, w, h, dtype);
    if (rc) return rc;
    if (!dst || pd < (size_t)w * dtype) return -3;
    if (conn != 4 && conn != 8) return -4;
    if ((dir != 1 && dir != -1) || bw <= 0 || bh <= 0) return -6;
    if (dtype == 1) pass_y_u8(src, dst, mask, ps, pd, pi, w, h, conn, dir, bw, bh);
    else pass_y_u16(src, dst, mask, ps, pd, pi, w, h, conn, dir, bw, bh);
    return 0;
}

int oracle_pass_raster(const void* src, void* dst, const void* mask, size_t ps, size_t pd,
                       size_t pi, int w, int h, int dtype, int conn, int dir, int bw, int bh) {
    int rc = check_args(src, mask, ps, pi, w, h, dtype);
    if (rc) return rc;
    if (!dst || pd < (size_t)w * dtype) return -3;
    if (conn != 4 && conn != 8) return -4;
    if ((dir != 1 && dir != -1) || bw <= 0 || bh <= 0) return -6;
    if (dtype == 1) pass_raster_u8(src, dst, mask, ps, pd, pi, w, h, conn, dir, bw, bh);
    else pass_raster_u16(src, dst, mask, ps, pd, pi, w, h, conn, dir, bw, bh);
    return 0;
}

long oracle_unstable_blocks(const void* marker, const void* mask, size_t pj, size_t pi, int w,
                            int h, int dtype, int conn, int mw, int mh, uint8_t* flags) {
    int rc = check_args(marker, mask, pj, pi, w, h, dtype);
    if (rc) return rc;
    if (conn != 4 && conn != 8) return -4;
    if (mw <= 0 || mh <= 0 || !flags) return -6;
    return dtype == 1 ? unstable_u8(marker, mask, pj, pi, w, h, conn, mw, mh, flags)
                      : unstable_u16(marker, mask, pj, pi, w, h, conn, mw, mh, flags);
}
