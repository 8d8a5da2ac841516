// Synthetic workload generators (host C++, C ABI declared in include/iwpp_synth.h).
// Integer-only and deterministic: a seed names one image on every x86-64 host.
#include "iwpp_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// xoshiro256** (Blackman & Vigna), seeded through splitmix64.
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // Uniform integer in [0, n) by the multiply-shift method.
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    // Uniform double in [0, 1) with 53 random bits.
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int hw_threads(int requested) {
    if (requested > 0) return requested;
    unsigned n = std::thread::hardware_concurrency();
    return n ? (int)std::min(n, 64u) : 1;
}

template <class F>
void parallel_for(int n, int nthreads, F&& f) {
    nthreads = std::max(1, std::min(nthreads, n));
    if (nthreads == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    const int chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        ts.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& th : ts) th.join();
}

inline int reflect(int i, int n) {
    // scipy-style 'reflect' (d c b a | a b c d | d c b a), repeated for tiny n.
    while (i < 0 || i >= n) {
        if (i < 0) i = -i - 1;
        if (i >= n) i = 2 * n - i - 1;
    }
    return i;
}

// Box widths for n=3 successive box filters approximating a Gaussian (Kovesi 2010).
void box_widths(double sigma, int out[3]) {
    const int n = 3;
    const double wideal = std::sqrt(12.0 * sigma * sigma / n + 1.0);
    int wl = (int)std::floor(wideal);
    if (wl % 2 == 0) wl--;
    if (wl < 1) wl = 1;
    const int wu = wl + 2;
    const double mideal =
        (12.0 * sigma * sigma - n * wl * wl - 4.0 * n * wl - 3.0 * n) / (-4.0 * wl - 4.0);
    const long m = std::lround(mideal);
    for (int i = 0; i < n; ++i) out[i] = i < m ? wl : wu;
}

// One horizontal box pass of odd width w over a W x H uint16 plane.
void box_h(const uint16_t* in, uint16_t* out, int W, int H, int w, int nth) {
    const int r = w / 2;
    parallel_for(H, nth, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y) {
            const uint16_t* src = in + (size_t)y * W;
            uint16_t* dst = out + (size_t)y * W;
            uint32_t sum = 0;
            for (int k = -r; k <= r; ++k) sum += src[reflect(k, W)];
            for (int x = 0; x < W; ++x) {
                dst[x] = (uint16_t)((sum + (uint32_t)w / 2) / (uint32_t)w);
                sum += src[reflect(x + r + 1, W)];
                sum -= src[reflect(x - r, W)];
            }
        }
    });
}

// One vertical box pass, columns processed in strips of 64 for cache reuse.
void box_v(const uint16_t* in, uint16_t* out, int W, int H, int w, int nth) {
    const int r = w / 2;
    const int nstrips = (W + 63) / 64;
    parallel_for(nstrips, nth, [&](int s0, int s1) {
        std::vector<uint32_t> sum(64);
        for (int s = s0; s < s1; ++s) {
            const int x0 = s * 64, x1 = std::min(W, x0 + 64), nx = x1 - x0;
            std::fill(sum.begin(), sum.end(), 0u);
            for (int k = -r; k <= r; ++k) {
                const uint16_t* row = in + (size_t)reflect(k, H) * W + x0;
                for (int i = 0; i < nx; ++i) sum[i] += row[i];
            }
            for (int y = 0; y < H; ++y) {
                uint16_t* dst = out + (size_t)y * W + x0;
                for (int i = 0; i < nx; ++i)
                    dst[i] = (uint16_t)((sum[i] + (uint32_t)w / 2) / (uint32_t)w);
                const uint16_t* add = in + (size_t)reflect(y + r + 1, H) * W + x0;
                const uint16_t* sub = in + (size_t)reflect(y - r, H) * W + x0;
                for (int i = 0; i < nx; ++i) sum[i] += add[i] - sub[i];
            }
        }
    });
}

template <class T>
T* row_ptr(void* base, size_t pitch, int y) {
    return (T*)((uint8_t*)base + (size_t)y * pitch);
}
template <class T>
const T* row_ptr(const void* base, size_t pitch, int y) {
    return (const T*)((const uint8_t*)base + (size_t)y * pitch);
}

template <class T>
int synth_mask_t(void* mask, size_t pitch, int W, int H, int maxval, double sigma, int ndisks,
                 int rmin, int rmax, uint64_t seed, int nth) {
    const size_t N = (size_t)W * H;
    std::vector<uint16_t> a(N), b(N);
    // Noise: 8 bytes per splitmix64 hash of (seed, row, x/8); kept with 8 fractional bits.
    const uint64_t rowkey = splitmix64(seed ^ 0x6E6F697365ull);
    parallel_for(H, nth, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y) {
            uint16_t* row = a.data() + (size_t)y * W;
            for (int x8 = 0; x8 < W; x8 += 8) {
                uint64_t h = splitmix64(rowkey + (uint64_t)y * 0x100000001B3ull + (uint64_t)(x8 / 8));
                for (int k = 0; k < 8 && x8 + k < W; ++k)
                    row[x8 + k] = (uint16_t)(((h >> (8 * k)) & 0xFF) << 8);
            }
        }
    });
    if (sigma > 0) {
        int wd[3];
        box_widths(sigma, wd);
        box_h(a.data(), b.data(), W, H, wd[0], nth);
        box_h(b.data(), a.data(), W, H, wd[1], nth);
        box_h(a.data(), b.data(), W, H, wd[2], nth);
        box_v(b.data(), a.data(), W, H, wd[0], nth);
        box_v(a.data(), b.data(), W, H, wd[1], nth);
        box_v(b.data(), a.data(), W, H, wd[2], nth);
    }
    uint16_t lo = 0xFFFF, hi = 0;
    for (size_t i = 0; i < N; ++i) {
        lo = std::min(lo, a[i]);
        hi = std::max(hi, a[i]);
    }
    const uint64_t d = (uint64_t)(hi - lo);
    parallel_for(H, nth, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y) {
            T* out = row_ptr<T>(mask, pitch, y);
            const uint16_t* in = a.data() + (size_t)y * W;
            for (int x = 0; x < W; ++x) {
                uint64_t v = 0;
                if (d) {  // round((in - lo) * maxval / d), ties to even
                    const uint64_t q = (uint64_t)(in[x] - lo) * (uint64_t)maxval;
                    v = q / d;
                    const uint64_t rem = q % d;
                    if (2 * rem > d || (2 * rem == d && (v & 1))) ++v;
                }
                out[x] = (T)v;
            }
        }
    });
    Xoshiro rng(splitmix64(seed ^ 0x6469736B73ull));
    for (int k = 0; k < ndisks; ++k) {
        const int cx = (int)rng.below((uint64_t)W);
        const int cy = (int)rng.below((uint64_t)H);
        const int r = rmin + (int)rng.below((uint64_t)(rmax - rmin + 1));
        const T val = (T)rng.below((uint64_t)maxval + 1);
        for (int dy = -r; dy <= r; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= H) continue;
            const int64_t rr = (int64_t)r * r - (int64_t)dy * dy;
            int dx = (int)std::sqrt((double)rr);
            while ((int64_t)(dx + 1) * (dx + 1) <= rr) ++dx;
            while ((int64_t)dx * dx > rr) --dx;
            T* out = row_ptr<T>(mask, pitch, y);
            const int xa = std::max(0, cx - dx), xb = std::min(W - 1, cx + dx);
            for (int x = xa; x <= xb; ++x) out[x] = std::max(out[x], val);
        }
    }
    return 0;
}

template <class T>
void hdome_t(void* marker, const void* mask, size_t pj, size_t pi, int W, int H, int h) {
    for (int y = 0; y < H; ++y) {
        T* J = row_ptr<T>(marker, pj, y);
        const T* I = row_ptr<T>(mask, pi, y);
        for (int x = 0; x < W; ++x) J[x] = (T)(I[x] > h ? I[x] - h : 0);
    }
}

template <class T>
void seeds_t(void* marker, const void* mask, size_t pj, size_t pi, int W, int H, long count,
             uint64_t seed) {
    for (int y = 0; y < H; ++y) std::memset(row_ptr<T>(marker, pj, y), 0, (size_t)W * sizeof(T));
    const uint64_t N = (uint64_t)W * H;
    std::vector<uint64_t> bits((N + 63) / 64, 0);
    Xoshiro rng(splitmix64(seed ^ 0x7365656473ull));
    long placed = 0;
    while (placed < count) {
        const uint64_t i = rng.below(N);
        if (bits[i / 64] >> (i % 64) & 1) continue;
        bits[i / 64] |= 1ull << (i % 64);
        const int y = (int)(i / (uint64_t)W), x = (int)(i % (uint64_t)W);
        row_ptr<T>(marker, pj, y)[x] = row_ptr<T>(mask, pi, y)[x];
        ++placed;
    }
}

// uint16 images need 2-byte aligned rows (pixels are accessed as uint16_t).
bool bad_image(const void* a, size_t pitch, int W, int H, int dtype) {
    return !a || W <= 0 || H <= 0 || (dtype != 1 && dtype != 2) || pitch < (size_t)W * dtype ||
           pitch % dtype || (uintptr_t)a % dtype;
}

}  // namespace

extern "C" int iwpp_synth_mask(void* mask, size_t pitch, int W, int H, int dtype, int maxval,
                               double sigma, int ndisks, int rmin, int rmax, uint64_t seed,
                               int nthreads) {
    if (bad_image(mask, pitch, W, H, dtype)) return -1;
    if (maxval <= 0 || maxval > (dtype == 1 ? 255 : 65535)) return -2;
    if (ndisks < 0 || (ndisks > 0 && (rmin < 0 || rmax < rmin))) return -3;
    const int nth = hw_threads(nthreads);
    return dtype == 1 ? synth_mask_t<uint8_t>(mask, pitch, W, H, maxval, sigma, ndisks, rmin,
                                              rmax, seed, nth)
                      : synth_mask_t<uint16_t>(mask, pitch, W, H, maxval, sigma, ndisks, rmin,
                                               rmax, seed, nth);
}

extern "C" int iwpp_synth_marker_hdome(void* marker, const void* mask, size_t pj, size_t pi,
                                       int W, int H, int dtype, int h) {
    if (bad_image(marker, pj, W, H, dtype) || bad_image(mask, pi, W, H, dtype) || h < 0)
        return -1;
    if (dtype == 1) hdome_t<uint8_t>(marker, mask, pj, pi, W, H, h);
    else hdome_t<uint16_t>(marker, mask, pj, pi, W, H, h);
    return 0;
}

extern "C" int iwpp_synth_marker_seeds(void* marker, const void* mask, size_t pj, size_t pi,
                                       int W, int H, int dtype, long count, uint64_t seed) {
    if (bad_image(marker, pj, W, H, dtype) || bad_image(mask, pi, W, H, dtype)) return -1;
    if (count < 0 || (uint64_t)count > (uint64_t)W * H) return -2;
    if (dtype == 1) seeds_t<uint8_t>(marker, mask, pj, pi, W, H, count, seed);
    else seeds_t<uint16_t>(marker, mask, pj, pi, W, H, count, seed);
    return 0;
}

extern "C" int iwpp_synth_batch_shapes(int n, uint64_t seed, int* widths, int* heights,
                                       int* kinds, int* hs) {
    if (n <= 0 || !widths || !heights || !kinds || !hs) return -1;
    Xoshiro rng(splitmix64(seed ^ 0x6261746368ull));
    for (int i = 0; i < n; ++i) {
        const double u = rng.unit();
        int s = (int)std::floor(256.0 * std::pow(16.0, u));
        s = std::min(4096, std::max(256, s));
        const double a = 1.0 + 3.0 * rng.unit();
        int w = s, h = std::max(1, (int)std::floor(s / a));
        if ((i / 2) % 2 == 1) std::swap(w, h);
        widths[i] = w;
        heights[i] = h;
        kinds[i] = i % 2;
        hs[i] = 10 + (int)rng.below(51);
    }
    return 0;
}

extern "C" int iwpp_synth_random_case(void* marker, void* mask, size_t pj, size_t pi, int W,
                                      int H, uint64_t seed) {
    if (bad_image(marker, pj, W, H, 1) || bad_image(mask, pi, W, H, 1)) return -1;
    for (int y = 0; y < H; ++y) {
        uint8_t* J = row_ptr<uint8_t>(marker, pj, y);
        uint8_t* I = row_ptr<uint8_t>(mask, pi, y);
        for (int x = 0; x < W; ++x) {
            const uint64_t h = splitmix64(seed * 0x100000001B3ull + (uint64_t)y * W + x);
            const int m = (int)(h & 0xFF);
            const int r = (int)((h >> 8) & 0xFF) % (256 - m);
            J[x] = (uint8_t)m;
            I[x] = (uint8_t)(m + r);
        }
    }
    return 0;
}
