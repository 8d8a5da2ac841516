// PGM reader / writer (include/iwpp_io.h).  Host code.
#include "iwpp_io.h"

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

// Next header token (whitespace separated, '#' comments skipped).  Returns
// false at end of file or when the token is longer than the buffer.
bool token(FILE* f, char* buf, size_t n) {
    int c = fgetc(f);
    while (c != EOF) {
        if (c == '#') {
            while (c != EOF && c != '\n' && c != '\r') c = fgetc(f);
        } else if (isspace(c)) {
            c = fgetc(f);
        } else {
            break;
        }
    }
    if (c == EOF) return false;
    size_t k = 0;
    while (c != EOF && !isspace(c) && c != '#') {
        if (k + 1 >= n) return false;
        buf[k++] = (char)c;
        c = fgetc(f);
    }
    buf[k] = 0;
    if (c == '#') ungetc(c, f);  // a comment right after the token
    // P5: exactly one whitespace character separates maxval from the data,
    // already consumed above.
    return true;
}

bool number(FILE* f, long lo, long hi, long* out) {
    char buf[32];
    if (!token(f, buf, sizeof buf)) return false;
    char* end = nullptr;
    const long v = strtol(buf, &end, 10);
    if (!end || *end || v < lo || v > hi) return false;
    *out = v;
    return true;
}

}  // namespace

extern "C" int iwpp_pgm_read(const char* path, void** data, int* width, int* height,
                             int* maxval) {
    if (!path || !data || !width || !height || !maxval) return IWPP_IO_ERR_ARG;
    *data = nullptr;
    File F;
    F.f = fopen(path, "rb");
    if (!F.f) return IWPP_IO_ERR_OPEN;
    char magic[8];
    if (!token(F.f, magic, sizeof magic)) return IWPP_IO_ERR_FORMAT;
    const bool binary = strcmp(magic, "P5") == 0;
    if (!binary && strcmp(magic, "P2") != 0) return IWPP_IO_ERR_FORMAT;
    long w, h, mv;
    if (!number(F.f, 1, 1L << 30, &w) || !number(F.f, 1, 1L << 30, &h) ||
        !number(F.f, 1, 65535, &mv))
        return IWPP_IO_ERR_FORMAT;
    const size_t bpp = mv > 255 ? 2 : 1;
    const size_t n = (size_t)w * (size_t)h;
    if (n / (size_t)w != (size_t)h) return IWPP_IO_ERR_FORMAT;
    uint8_t* buf = (uint8_t*)malloc(n * bpp);
    if (!buf) return IWPP_IO_ERR_NOMEM;
    bool ok = true;
    if (binary) {
        ok = fread(buf, bpp, n, F.f) == n;
        if (ok && bpp == 2) {  // big-endian in the file
            uint16_t* p = (uint16_t*)buf;
            for (size_t i = 0; i < n; ++i) {
                const uint8_t* b = buf + 2 * i;
                p[i] = (uint16_t)((b[0] << 8) | b[1]);
            }
        }
        if (ok) {
            for (size_t i = 0; i < n && ok; ++i) {
                const long v = bpp == 2 ? ((uint16_t*)buf)[i] : buf[i];
                ok = v <= mv;
            }
        }
    } else {
        for (size_t i = 0; i < n && ok; ++i) {
            long v;
            ok = number(F.f, 0, mv, &v);
            if (ok) {
                if (bpp == 2) ((uint16_t*)buf)[i] = (uint16_t)v;
                else buf[i] = (uint8_t)v;
            }
        }
    }
    if (!ok) {
        free(buf);
        return IWPP_IO_ERR_FORMAT;
    }
    *data = buf;
    *width = (int)w;
    *height = (int)h;
    *maxval = (int)mv;
    return IWPP_IO_OK;
}

extern "C" int iwpp_pgm_write(const char* path, const void* data, size_t pitch, int width,
                              int height, int maxval) {
    if (!path || !data || width <= 0 || height <= 0 || maxval < 1 || maxval > 65535)
        return IWPP_IO_ERR_ARG;
    const size_t bpp = maxval > 255 ? 2 : 1;
    if (pitch < (size_t)width * bpp || pitch % bpp || (uintptr_t)data % bpp) return IWPP_IO_ERR_ARG;
    File F;
    F.f = fopen(path, "wb");
    if (!F.f) return IWPP_IO_ERR_OPEN;
    if (fprintf(F.f, "P5\n%d %d\n%d\n", width, height, maxval) < 0) return IWPP_IO_ERR_OPEN;
    uint8_t* row = (uint8_t*)malloc((size_t)width * bpp);
    if (!row) return IWPP_IO_ERR_NOMEM;
    bool ok = true;
    for (int y = 0; y < height && ok; ++y) {
        const uint8_t* src = (const uint8_t*)data + (size_t)y * pitch;
        if (bpp == 2) {
            for (int x = 0; x < width; ++x) {
                const uint16_t v = ((const uint16_t*)src)[x];
                row[2 * x] = (uint8_t)(v >> 8);
                row[2 * x + 1] = (uint8_t)(v & 0xFF);
            }
        } else {
            memcpy(row, src, (size_t)width);
        }
        ok = fwrite(row, bpp, (size_t)width, F.f) == (size_t)width;
    }
    free(row);
    if (!ok || fflush(F.f) != 0) return IWPP_IO_ERR_OPEN;
    return IWPP_IO_OK;
}

extern "C" void iwpp_pgm_free(void* data) { free(data); }
