// Learned schedule: per-tile features, and an MLP 16-64-64-2 (ReLU) that maps
// them to (sweep rounds before the queue phase, tile priority).  Inference
// uses BF16 weights with FP32 accumulation; training keeps FP32 master
// weights, runs the forward pass with the BF16-rounded weights and applies
// Adam.  The schedule only chooses which tiles the sweep rounds touch and
// the order in which megapixels enter the queue, so reconstruction results
// do not depend on it.
#include "iwpp_internal.h"
#include "iwpp_learned.h"

namespace iwpp {

#define DEV static __device__ __forceinline__

// ------------------------------------------------------------- features ----
// Raw per-tile statistics over every `sub`-th row of the tile (from its first):
//  0 sum I, 1 sum J, 2 sum (I-J), 3 sum (I-J)^2, 4 max(0, max (I-J)),
//  5 count J==I, 6 count J==0, 7 count seeds (J==I, J>0), 8 max J,
//  9 sum |I(x+1)-I(x)| over pixel pairs (x even, x+1 in the tile),
//  10..13 border inflow top/bottom/left/right: sum of max(0, min(J(q), I(p))
//  - J(p)) over border pixels p, q the pixel across the border: every pixel
//  of the top and bottom rows, the left and right columns on the sampled rows,
//  14 pixels sampled, 15 max I.
// Work split: one workgroup per (tile, slab of slab_rows(dtype) sampled rows); partial
// statistics go to part[tile][slab][16]; a second kernel combines and
// normalises them per image.
// Rows are read as VB-byte vectors (16 when rows are 16-byte aligned).
// Sampled rows per workgroup: 8 for uint8 (the integer statistics leave a
// workgroup bound by its load and reduction latency: 2048 workgroups of 4 rows
// took 28 us at 32768^2, 1024 of 8 22 us), 4 for uint16 (16 us against 19.5
// with 8 at 16384^2).
constexpr int slab_rows(int dtype) { return dtype == 1 ? 8 : 4; }
constexpr int kRawLds = 256;  // tiles of one unit whose statistics the plan keeps in LDS
constexpr int kPlanThreads = 512;  // workgroup of the plan's network: 8 waves, 256 VGPRs each
constexpr int kPlanTiles = kPlanThreads / 64;  // tiles per workgroup (one image): one per wave
static_assert(kOrderMax <= kPlanThreads, "the seeding order takes one tile per thread");
constexpr int kStats = 16;

DEV bool is_max_stat(int k) { return k == 4 || k == 8 || k == 15; }

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
DEV u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
// Byte b of w, for b in {0, 2} (lo) or {1, 3} (hi), as the two 16-bit halves.
DEV u16x2 bytes_lo(uint32_t w) { return as_u16x2(__builtin_amdgcn_perm(0u, w, 0x0c020c00u)); }
DEV u16x2 bytes_hi(uint32_t w) { return as_u16x2(__builtin_amdgcn_perm(0u, w, 0x0c030c01u)); }
// High bit of every zero byte of w (exact: the low seven bits plus 0x7F
// carry into the high bit, never into the next byte).
DEV uint32_t zero_bytes(uint32_t w) { return ~(((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u; }

DEV float wave_reduce(float v, bool mx) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float o = __shfl_xor(v, 32 >> i, 64);
        v = mx ? fmaxf(v, o) : v + o;
    }
    return v;
}

// Tiles of several images are numbered consecutively: image i owns global
// tiles tile_base[i] .. tile_base[i+1]-1 (tile_base == nullptr: one image).
DEV int image_of_tile(const uint32_t* tile_base, int n, int gt) {
    if (!tile_base) return 0;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)tile_base[mid] <= gt) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class T, int VB>
__global__ __launch_bounds__(256) void k_tile_features_part(Descs ds, const uint32_t* tile_base,
                                                             int tile, int sub, int nslab,
                                                             float* part) {
    constexpr int NPX = VB / (int)sizeof(T);
    const int gt = blockIdx.x / nslab, sl = blockIdx.x % nslab;
    const int img = image_of_tile(tile_base, ds.n, gt);
    const ImgDesc dd = ds.d ? ds.d[img] : ds.one;
    const Image im{dd.J, dd.I, dd.pj, dd.pi, dd.W, dd.H};
    const int ntx = (im.W + tile - 1) / tile;
    const int t = gt - (tile_base ? (int)tile_base[img] : 0);
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * tile, y0 = ty * tile;
    const int x1 = min(im.W, x0 + tile), y1 = min(im.H, y0 + tile);
    const int bx0 = x0 * (int)sizeof(T), bx1 = x1 * (int)sizeof(T);
    const int rowbytes = im.W * (int)sizeof(T);
    const int nvec = (bx1 - bx0 + VB - 1) / VB;
    float acc[kStats];
#pragma unroll
    for (int k = 0; k < kStats; ++k) acc[k] = 0.f;
    // The statistics of a 32-bit word's pixels at a time in integer registers
    // (sums of absolute differences, dot products or 16-bit squares, zero
    // counts, maxima on 16-bit halves): exact sums, converted to float once.
    // uint8: a thread's sums stay below 2^24, so they are the integers float
    // accumulation pixel by pixel would hold; uint16: the squared gaps are
    // summed in 64 bits (float accumulation rounded them past 2^24).
    uint32_t s_c = 0, s_j = 0, s_cc = 0, s_jj = 0, s_cj = 0, s_pair = 0;
    uint32_t n_eq = 0, n_jz = 0, n_seed = 0, n_px = 0;
    unsigned long long s_g2 = 0;
    u16x2 mx_c = {0, 0}, mx_j = {0, 0}, mx_g = {0, 0};
    constexpr int kSlab = slab_rows((int)sizeof(T));
    const int items = kSlab * nvec;
    // A thread's items (at most IPT: tile <= 2048 pixels of <= 2 bytes, 16-byte
    // vectors, 8 rows, 256 threads) are all loaded before any is used, from
    // addresses clamped into the row, the bytes past the row cleared after:
    // no branch around a load, so the loads are in flight together.
    constexpr int IPT = kSlab * (2048 * (int)sizeof(T) / VB) / 256;  // 2048-pixel tiles in one go
    for (int base = 0; base < items; base += IPT * (int)blockDim.x) {  // once for tiles <= 2048 px
    uint32_t jv[IPT][VB / 4], iv[IPT][VB / 4];
    int ov[IPT];
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        const int i = base + threadIdx.x + u * (int)blockDim.x;
        const int r = i / max(nvec, 1), v = i - r * max(nvec, 1);
        const int y = min(y0 + (sl * kSlab + r) * sub, im.H - 1);
        const int o = bx0 + v * VB;  // VB-aligned (tiles start at multiples of 64 px)
        const bool live = i < items && y0 + (sl * kSlab + r) * sub < y1;
        ov[u] = live ? o : -1;
        const int oc = (live && o < rowbytes) ? o : 0;
        const uint8_t* jr = im.J + (size_t)y * im.pj + oc;
        const uint8_t* ir = im.I + (size_t)y * im.pi + oc;
        if constexpr (VB == 16) {
            const uint4 a = *reinterpret_cast<const uint4*>(jr);
            const uint4 b = *reinterpret_cast<const uint4*>(ir);
            jv[u][0] = a.x; jv[u][1] = a.y; jv[u][2] = a.z; jv[u][3] = a.w;
            iv[u][0] = b.x; iv[u][1] = b.y; iv[u][2] = b.z; iv[u][3] = b.w;
        } else {
            jv[u][0] = *reinterpret_cast<const uint32_t*>(jr);
            iv[u][0] = *reinterpret_cast<const uint32_t*>(ir);
        }
    }
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        const int o = ov[u];
        if (o < 0) continue;
        uint32_t jw[VB / 4], iw[VB / 4];
        const int valid = rowbytes - o;  // bytes of this vector inside the row
#pragma unroll
        for (int w = 0; w < VB / 4; ++w) {
            const int vb = valid - 4 * w;
            const uint32_t keep = vb >= 4 ? 0xFFFFFFFFu : vb <= 0 ? 0u : (0xFFFFFFFFu >> (8 * (4 - vb)));
            jw[w] = jv[u][w] & keep;
            iw[w] = iv[u][w] & keep;
        }
        const int xv = o / (int)sizeof(T);  // first pixel of the vector
        if constexpr (sizeof(T) == 1) {
            const int nv = min(NPX, x1 - xv);  // pixels of the vector inside the tile
            n_px += (uint32_t)nv;
#pragma unroll
            for (int w = 0; w < VB / 4; ++w) {
                const int vb = min(max(nv - 4 * w, 0), 4);
                const uint32_t bm = vb >= 4 ? 0xFFFFFFFFu : (1u << (8 * vb)) - 1u;
                const uint32_t c = iw[w] & bm, j = jw[w] & bm;
                s_c = __builtin_amdgcn_sad_u8(c, 0u, s_c);
                s_j = __builtin_amdgcn_sad_u8(j, 0u, s_j);
                s_cc = __builtin_amdgcn_udot4(c, c, s_cc, false);
                s_jj = __builtin_amdgcn_udot4(j, j, s_jj, false);
                s_cj = __builtin_amdgcn_udot4(c, j, s_cj, false);
                // pairs (0, 1) and (2, 3) whose second pixel is in the tile
                const uint32_t pm = (vb >= 2 ? 0x000000FFu : 0u) | (vb >= 4 ? 0x00FF0000u : 0u);
                s_pair = __builtin_amdgcn_sad_u8(c & pm, (c >> 8) & pm, s_pair);
                const uint32_t hv = bm & 0x80808080u;
                const uint32_t zx = zero_bytes(c ^ j) & hv, zj = zero_bytes(j) & hv;
                n_eq += (uint32_t)__builtin_popcount(zx);
                n_jz += (uint32_t)__builtin_popcount(zj);
                n_seed += (uint32_t)__builtin_popcount(zx & ~zj);
                const u16x2 cl = bytes_lo(c), ch = bytes_hi(c), jl = bytes_lo(j), jh = bytes_hi(j);
                mx_c = __builtin_elementwise_max(mx_c, __builtin_elementwise_max(cl, ch));
                mx_j = __builtin_elementwise_max(mx_j, __builtin_elementwise_max(jl, jh));
                mx_g = __builtin_elementwise_max(
                    mx_g, __builtin_elementwise_max(__builtin_elementwise_sub_sat(cl, jl),
                                                    __builtin_elementwise_sub_sat(ch, jh)));
            }
        } else {
            const int nv = min(NPX, x1 - xv);  // pixels of the vector inside the tile
            n_px += (uint32_t)nv;
#pragma unroll
            for (int w = 0; w < VB / 4; ++w) {
                const int vp = min(max(nv - 2 * w, 0), 2);  // pixels of the word in the tile
                const uint32_t bm = vp >= 2 ? 0xFFFFFFFFu : vp == 1 ? 0xFFFFu : 0u;
                const uint32_t c = iw[w] & bm, j = jw[w] & bm;
                s_c = __builtin_amdgcn_sad_u16(c, 0u, s_c);
                s_j = __builtin_amdgcn_sad_u16(j, 0u, s_j);
                const u16x2 cp = as_u16x2(c), jp = as_u16x2(j);
                const u16x2 d = __builtin_elementwise_max(cp, jp) - __builtin_elementwise_min(cp, jp);
                s_g2 += (unsigned long long)((uint32_t)d.x * d.x) + (uint32_t)d.y * d.y;
                // the pair (first, second pixel of the word) when both are in the tile
                if (vp >= 2) s_pair += (uint32_t)max(cp.x, cp.y) - min(cp.x, cp.y);
                const uint32_t x = c ^ j;
                const bool e0 = vp >= 1 && (x & 0xFFFFu) == 0u, e1 = vp >= 2 && (x >> 16) == 0u;
                const bool z0 = vp >= 1 && (j & 0xFFFFu) == 0u, z1 = vp >= 2 && (j >> 16) == 0u;
                n_eq += (uint32_t)e0 + (uint32_t)e1;
                n_jz += (uint32_t)z0 + (uint32_t)z1;
                n_seed += (uint32_t)(e0 && !z0) + (uint32_t)(e1 && !z1);
                mx_c = __builtin_elementwise_max(mx_c, cp);
                mx_j = __builtin_elementwise_max(mx_j, jp);
                mx_g = __builtin_elementwise_max(mx_g, __builtin_elementwise_sub_sat(cp, jp));
            }
        }
    }
    }
    {
        acc[0] = (float)s_c;
        acc[1] = (float)s_j;
        acc[2] = (float)((long long)s_c - (long long)s_j);
        if constexpr (sizeof(T) == 1) acc[3] = (float)(s_cc + s_jj - 2u * s_cj);  // sum (c - j)^2
        else acc[3] = (float)s_g2;
        acc[4] = (float)max(mx_g.x, mx_g.y);
        acc[5] = (float)n_eq;
        acc[6] = (float)n_jz;
        acc[7] = (float)n_seed;
        acc[8] = (float)max(mx_j.x, mx_j.y);
        acc[9] = (float)s_pair;
        acc[14] = (float)n_px;
        acc[15] = (float)max(mx_c.x, mx_c.y);
    }
    // border inflow: rows y0 / y1-1 in the first slab, the left and right
    // columns on this slab's sampled rows (a column read touches one cache
    // line per row: on every row it cost more than the rest of the kernel)
    auto J = [&](int x, int y) -> float {
        return (float)*reinterpret_cast<const T*>(im.J + (size_t)y * im.pj + (size_t)x * sizeof(T));
    };
    auto I = [&](int x, int y) -> float {
        return (float)*reinterpret_cast<const T*>(im.I + (size_t)y * im.pi + (size_t)x * sizeof(T));
    };
    if (sl == 0) {  // groups of 8 columns per thread, loads (clamped) before use
        const int ya = max(y0 - 1, 0), yb = min(y1, im.H - 1);
        for (int xb = x0; xb < x1; xb += 8 * (int)blockDim.x) {
            float a[8][6];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int x = min(xb + (int)threadIdx.x + u * (int)blockDim.x, x1 - 1);
                a[u][0] = J(x, ya);
                a[u][1] = I(x, y0);
                a[u][2] = J(x, y0);
                a[u][3] = J(x, yb);
                a[u][4] = I(x, y1 - 1);
                a[u][5] = J(x, y1 - 1);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (xb + (int)threadIdx.x + u * (int)blockDim.x >= x1) continue;
                if (y0 > 0) acc[10] += fmaxf(0.f, fminf(a[u][0], a[u][1]) - a[u][2]);
                if (y1 < im.H) acc[11] += fmaxf(0.f, fminf(a[u][3], a[u][4]) - a[u][5]);
            }
        }
    }
    if ((int)threadIdx.x < kSlab) {
        const int y = y0 + (sl * kSlab + (int)threadIdx.x) * sub;
        if (y < y1) {
            if (x0 > 0) acc[12] += fmaxf(0.f, fminf(J(x0 - 1, y), I(x0, y)) - J(x0, y));
            if (x1 < im.W) acc[13] += fmaxf(0.f, fminf(J(x1, y), I(x1 - 1, y)) - J(x1 - 1, y));
        }
    }
    __shared__ float red[kStats][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kStats; ++k) {
        const float v = wave_reduce(acc[k], is_max_stat(k));
        if (lane == 0) red[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kStats) {
        const int k = threadIdx.x;
        const bool mx = is_max_stat(k);
        float v = red[k][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = mx ? fmaxf(v, red[k][w]) : v + red[k][w];
        part[((size_t)gt * nslab + sl) * kStats + k] = v;
    }
}

// Normalised features from raw statistics.  M = max mask value over all tiles
// (0 -> 1); tile_px = sampled pixels of a full tile.  Features 1 and 14 carry
// the connectivity (8 -> 1, 4 -> 0) and the pixel type (uint16 -> 1, uint8 -> 0).
DEV void normalise(const float* r, float M, int tile, int sub, int dtype, int conn, float log2px,
                   float* f) {
    const float n = fmaxf(r[14], 1.f);
    const float inv = 1.f / fmaxf(M, 1.f);
    f[0] = r[0] / n * inv;
    f[1] = conn == 8 ? 1.f : 0.f;
    f[2] = r[2] / n * inv;
    f[3] = sqrtf(r[3] / n) * inv;
    f[4] = r[4] * inv;
    f[5] = r[5] / n;
    f[6] = r[6] / n;
    f[7] = r[7] / n * 100.f;
    f[8] = r[8] * inv;
    f[9] = r[9] / n * inv * 10.f;
    const float side = (float)tile;
    f[10] = r[10] / side * inv;
    f[11] = r[11] / side * inv;
    f[12] = r[12] * (float)sub / side * inv;  // columns sampled on every sub-th row
    f[13] = r[13] * (float)sub / side * inv;
    f[14] = dtype == 2 ? 1.f : 0.f;
    // the image's size (what a size rule would key on), at most 2^28 pixels:
    // larger images than the training set's largest are seen as that size
    f[15] = fminf(log2px, 28.f) / 32.f;
}

// Raw statistics of each tile from its slab partials: one wave per tile,
// lane = (slab group g = lane / 16, statistic k = lane % 16); each lane folds
// every fourth slab, then the four groups are folded across lanes (a fixed
// order: the result does not depend on scheduling).
// Writes the tile's 16 statistics to out[0..15].  A lane folds every fourth
// slab in order; its loads go out in groups of 8 before any is used.
DEV void fold_tile(const float* part, int nslab, int t, float* out) {
    const int lane = threadIdx.x & 63;
    const int k = lane & 15, g = lane >> 4;
    const bool mx = is_max_stat(k);
    float v = 0.f;  // identity of both folds (statistics are >= 0)
    constexpr int U = 8;
    for (int s0 = g; s0 < nslab; s0 += 4 * U) {
        float o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int sl = s0 + 4 * u;
            o[u] = sl < nslab ? part[((size_t)t * nslab + sl) * kStats + k] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v = mx ? fmaxf(v, o[u]) : v + o[u];
    }
    const float o16 = __shfl_xor(v, 16);
    v = mx ? fmaxf(v, o16) : v + o16;
    const float o32 = __shfl_xor(v, 32);
    v = mx ? fmaxf(v, o32) : v + o32;
    if (g == 0) out[k] = v;
}

__global__ __launch_bounds__(256) void k_features_tiles(const float* part, int nslab, int ntiles,
                                                         float* raw) {
    const int t = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (t >= ntiles) return;  // whole wave
    fold_tile(part, nslab, t, raw + (size_t)t * kStats);
}

// One workgroup per image: M = the image's largest mask value over its tiles'
// raw statistics, then its tiles' normalised features.
__global__ __launch_bounds__(256) void k_features_finish(Descs ds, const uint32_t* tile_base,
                                                          int ntiles, int tile, int sub, int dtype,
                                                          int conn, const float* raw, float* feat) {
    const ImgDesc dd = ds.d ? ds.d[blockIdx.x] : ds.one;
    const float log2px = log2f((float)dd.W * (float)dd.H);
    __shared__ float smax[256];
    const int t0 = tile_base ? (int)tile_base[blockIdx.x] : 0;
    const int t1 = tile_base ? (int)tile_base[blockIdx.x + 1] : ntiles;
    float m = 0.f;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x) m = fmaxf(m, raw[(size_t)t * kStats + 15]);
    smax[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        __syncthreads();
    }
    const float M = smax[0];
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x)
        normalise(raw + (size_t)t * 16, M, tile, sub, dtype, conn, log2px, feat + (size_t)t * 16);
}

// ------------------------------------------------------------------ MLP ----
// Parameter layout (kMlpParams floats): W1[64][16] b1[64] W2[64][64] b2[64]
// W3[2][64] b3[2].
constexpr int O_W1 = 0, O_B1 = O_W1 + 64 * 16, O_W2 = O_B1 + 64, O_B2 = O_W2 + 64 * 64,
              O_W3 = O_B2 + 64, O_B3 = O_W3 + 2 * 64;
static_assert(O_B3 + 2 == kMlpParams, "MLP layout");

// One wave per sample (grid-stride over samples); lane j owns hidden unit j
// of both layers and keeps its weights in registers (W1 row j, W2 row j, the
// two W3 entries of column j).  Inputs and layer-1 activations reach every
// lane through readlane; the outputs are wave sums.  FP32 accumulation.
// acts (optional, training): per sample h1[64] then h2[64] (after ReLU).
// A store to the host's pinned memory that no device cache holds.
DEV void st_host(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// Lane j's share of the weights: W1 row j, W2 row j, the two W3 entries of
// column j, and the biases.
struct MlpLane {
    float w1[16], w2[64], b1, b2, w30, w31, b30, b31;
};
// N consecutive weights from element `first` (16-byte aligned: the rows of W1
// and W2 start at multiples of 8 elements), as 16-byte loads.
template <bool BF16, int N>
DEV void load_row(const void* params, int first, float* w) {
    constexpr int PER = BF16 ? 8 : 4;  // weights per 16 bytes
    const uint4* p = reinterpret_cast<const uint4*>(
        reinterpret_cast<const char*>(params) + (size_t)first * (BF16 ? 2 : 4));
    uint4 v[N / PER];
#pragma unroll
    for (int i = 0; i < N / PER; ++i) v[i] = p[i];
#pragma unroll
    for (int i = 0; i < N / PER; ++i) {
        const uint32_t q[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if constexpr (BF16) {
                w[i * 8 + 2 * j] = __builtin_bit_cast(float, q[j] << 16);
                w[i * 8 + 2 * j + 1] = __builtin_bit_cast(float, q[j] & 0xFFFF0000u);
            } else {
                w[i * 4 + j] = __builtin_bit_cast(float, q[j]);
            }
        }
    }
}

template <bool BF16>
DEV MlpLane mlp_lane(const void* params, int lane) {
    auto W = [&](int i) -> float {
        return BF16 ? bf16_to_f32(reinterpret_cast<const uint16_t*>(params)[i])
                    : reinterpret_cast<const float*>(params)[i];
    };
    static_assert(O_W1 % 8 == 0 && O_W2 % 8 == 0, "weight rows 16-byte aligned");
    MlpLane m;
    load_row<BF16, 16>(params, O_W1 + lane * 16, m.w1);
    load_row<BF16, 64>(params, O_W2 + lane * 64, m.w2);
    m.b1 = W(O_B1 + lane);
    m.b2 = W(O_B2 + lane);
    m.w30 = W(O_W3 + lane);
    m.w31 = W(O_W3 + 64 + lane);
    m.b30 = W(O_B3);
    m.b31 = W(O_B3 + 1);
    return m;
}
// One sample: lane k < 16 holds input k in xv.  Returns both outputs in every
// lane, and the lane's hidden units h1, h2 (after ReLU).
DEV void mlp_sample(const MlpLane& m, float xv, float& y0, float& y1, float& h1, float& h2) {
    float a = m.b1;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        a = fmaf(m.w1[k], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv), k)), a);
    h1 = fmaxf(a, 0.f);
    a = m.b2;
#pragma unroll
    for (int k = 0; k < 64; ++k)
        a = fmaf(m.w2[k], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), k)), a);
    h2 = fmaxf(a, 0.f);
    y0 = m.w30 * h2;
    y1 = m.w31 * h2;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        y0 += __shfl_xor(y0, 32 >> i, 64);
        y1 += __shfl_xor(y1, 32 >> i, 64);
    }
    y0 += m.b30;
    y1 += m.b31;
}

template <bool BF16>
__global__ __launch_bounds__(256) void k_mlp_forward(const void* params, const float* x, int n,
                                                      float* y, float* acts) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    const MlpLane m = mlp_lane<BF16>(params, lane);
    for (int s = wave; s < n; s += nwaves) {
        const float xv = lane < 16 ? x[(size_t)s * 16 + lane] : 0.f;
        float y0, y1, h1, h2;
        mlp_sample(m, xv, y0, y1, h1, h2);
        if (acts) {
            acts[(size_t)s * 128 + lane] = h1;
            acts[(size_t)s * 128 + 64 + lane] = h2;
        }
        if (lane == 0) {
            y[(size_t)s * 2] = y0;
            y[(size_t)s * 2 + 1] = y1;
        }
    }
}

// The plan's network outputs from the slab partials.  A workgroup (8 waves)
// takes one unit: in a batch one image, for one image kPlanTiles of its
// tiles.  For each tile: fold its slabs (as k_features_tiles), normalise with
// the image's largest mask value (the maximum of statistic 15 over the
// image's partials: the same value k_features_finish takes from the folded
// tiles), run the network (BF16 weights, as k_mlp_forward) and write the two
// outputs to `out` -- the host's pinned staging, mapped.  The last workgroup
// to finish writes `seq` to `flag` (host memory as well), which the host
// polls: no copy and no event between the network and the host's plan.
// `count` (device, zero) counts finished workgroups and is left zero.
// With `order` (one image of at most kOrderMax tiles) the last workgroup also
// writes the seeding order: order[r] = the tile of rank r by decreasing
// priority (ties by tile index, NaN last), order[nt + r] = the first queue
// position of that tile's megapixels, order[2 nt] = their total; `prio`
// (device, nt floats) carries the priorities between workgroups.
__global__ __launch_bounds__(kPlanThreads) void k_plan_predict(Descs ds, const uint32_t* tile_base,
                                                        int ntiles, int nslab, int tile, int sub,
                                                        int dtype, int conn, const float* part,
                                                        float* raw, const void* params,
                                                        float* out, uint32_t* count,
                                                        uint32_t* flag, uint32_t seq,
                                                        float* prio, uint32_t* order) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const bool one = tile_base == nullptr;
    const int img = one ? 0 : (int)blockIdx.x;
    const ImgDesc dd = ds.d ? ds.d[img] : ds.one;
    const int i0 = one ? 0 : (int)tile_base[img], i1 = one ? ntiles : (int)tile_base[img + 1];
    const int t0 = one ? (int)blockIdx.x * kPlanTiles : i0;
    const int t1 = one ? min(ntiles, t0 + kPlanTiles) : i1;
    const MlpLane m = mlp_lane<true>(params, lane);  // in flight during the fold
    __shared__ float smax[16];
    __shared__ float key[kOrderMax];
    __shared__ uint32_t cnt[kOrderMax];
    __shared__ uint32_t wsum[16];
    __shared__ int last;
    // the image's largest mask value, from its partials
    float mx = 0.f;
    for (int i = i0 * nslab + (int)threadIdx.x; i < i1 * nslab; i += blockDim.x)
        mx = fmaxf(mx, part[(size_t)i * kStats + 15]);
    mx = wave_reduce(mx, true);
    if (lane == 0) smax[wv] = mx;
    const float log2px = log2f((float)dd.W * (float)dd.H);
    // This unit's tiles: fold, features and network.  rw holds their raw
    // statistics, y their outputs; both in LDS when they fit (instantiated
    // apart, so that the compiler sees LDS accesses).
    auto tiles = [&](float* rw, float* y) {
        for (int tb = t0 + wv; tb < t1; tb += 4 * nw) {  // four tiles' partials in flight
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = tb + u * nw;
                if (t < t1) fold_tile(part, nslab, t, rw + (size_t)(t - t0) * kStats);
            }
        }
        __syncthreads();
        float M = 0.f;
        for (int w = 0; w < nw; ++w) M = fmaxf(M, smax[w]);
        for (int t = t0 + wv; t < t1; t += nw) {
            float f[16];  // every lane computes them; lane k < 16 keeps feature k
            normalise(rw + (size_t)(t - t0) * kStats, M, tile, sub, dtype, conn, log2px, f);
            float xv = 0.f;
#pragma unroll
            for (int k = 0; k < 16; ++k) xv = lane == k ? f[k] : xv;
            float y0, y1, h1, h2;
            mlp_sample(m, xv, y0, y1, h1, h2);
            if (lane == 0) {
                y[(size_t)(t - t0) * 2] = y0;
                y[(size_t)(t - t0) * 2 + 1] = y1;
                if (order)  // shared with the other workgroups, past the caches
                    __hip_atomic_store(prio + t, isnan(y1) ? -INFINITY : y1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        for (int i = (int)threadIdx.x; i < 2 * (t1 - t0); i += blockDim.x)
            st_host(out + (size_t)t0 * 2 + i, y[i]);
    };
    if (t1 - t0 <= kRawLds) {
        __shared__ float sraw[kRawLds * kStats];
        __shared__ float sy[kRawLds * 2];
        tiles(sraw, sy);
    } else {
        tiles(raw + (size_t)t0 * kStats, raw + (size_t)ntiles * kStats + (size_t)t0 * 2);
    }
    // The outputs went out as stores that no device cache keeps (system scope
    // to the host, device scope for prio): waiting until they have completed
    // puts them before this workgroup's count, and so before the flag.  (A
    // release fence would do it too, but writes back the whole L2 -- dirty
    // with the row pass just before -- and cost ~10 us.)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t done = gridDim.x == 1 ? 1u
                                             : __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT) + 1u;
        last = done == gridDim.x;
        if (last && gridDim.x > 1) __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!last) return;
    if (order) {  // one image: every workgroup's priorities are in prio
        const int nt = ntiles, ntx = (dd.W + tile - 1) / tile, mpt = tile / 64;
        for (int t = (int)threadIdx.x; t < nt; t += blockDim.x)
            key[t] = __hip_atomic_load(prio + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        // rank of tile t = tiles before it in the order; megapixels by rank
        const int t = (int)threadIdx.x;  // nt <= kOrderMax = blockDim.x
        if (t < nt) {
            const float p = key[t];
            int rk = 0;
            for (int u = 0; u < nt; ++u) {
                const float q = key[u];
                rk += (q > p) || (q == p && u < t);
            }
            const int tx = t % ntx, ty = t / ntx;
            cnt[rk] = (uint32_t)(min(mpt, dd.nmx - tx * mpt) * min(mpt, dd.nmy - ty * mpt));
            order[rk] = (uint32_t)t;
        }
        __syncthreads();
        // exclusive prefix sum of cnt over ranks: lanes, then waves
        const uint32_t c = t < nt ? cnt[t] : 0u;
        uint32_t v = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(v, d, 64);
            if (lane >= d) v += o;
        }
        if (lane == 63) wsum[wv] = v;
        __syncthreads();
        uint32_t before = 0;
        for (int w = 0; w < wv; ++w) before += wsum[w];
        if (t < nt) order[nt + t] = before + v - c;
        if (t == nt - 1) order[2 * nt] = before + v;
    }
    if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Backward per sample for the weighted squared loss
//   L = sum_s sum_o lw[o] * (y[s][o] - t[s][o])^2 / (2 n)
// Writes d3[s][2], d2[s][64] (after ReLU'), d1[s][64].
template <bool BF16>
__global__ __launch_bounds__(256) void k_mlp_backward(const void* params, const float* acts,
                                                       const float* y, const float* t, int n,
                                                       float lw0, float lw1, float* d3, float* d2,
                                                       float* d1) {
    __shared__ float w[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x)
        w[i] = BF16 ? bf16_to_f32(reinterpret_cast<const uint16_t*>(params)[i])
                    : reinterpret_cast<const float*>(params)[i];
    __syncthreads();
    const float invn = 1.f / (float)n;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const float e0 = (y[(size_t)s * 2] - t[(size_t)s * 2]) * lw0 * invn;
        const float e1 = (y[(size_t)s * 2 + 1] - t[(size_t)s * 2 + 1]) * lw1 * invn;
        d3[(size_t)s * 2] = e0;
        d3[(size_t)s * 2 + 1] = e1;
        float g2[64];
#pragma unroll 8
        for (int k = 0; k < 64; ++k) {
            const float h2 = acts[(size_t)s * 128 + 64 + k];
            const float g = w[O_W3 + k] * e0 + w[O_W3 + 64 + k] * e1;
            g2[k] = h2 > 0.f ? g : 0.f;
            d2[(size_t)s * 64 + k] = g2[k];
        }
        for (int j = 0; j < 64; ++j) {
            float g = 0.f;
#pragma unroll
            for (int k = 0; k < 64; ++k) g = fmaf(w[O_W2 + k * 64 + j], g2[k], g);
            const float h1 = acts[(size_t)s * 128 + j];
            d1[(size_t)s * 64 + j] = h1 > 0.f ? g : 0.f;
        }
    }
}

// Parameter gradients: one thread per parameter, summing over samples.
__global__ __launch_bounds__(256) void k_mlp_grads(const float* x, const float* acts,
                                                    const float* d3, const float* d2,
                                                    const float* d1, int n, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float g = 0.f;
    if (p < O_B1) {
        const int o = p / 16, k = p % 16;
        for (int s = 0; s < n; ++s) g = fmaf(d1[(size_t)s * 64 + o], x[(size_t)s * 16 + k], g);
    } else if (p < O_W2) {
        const int o = p - O_B1;
        for (int s = 0; s < n; ++s) g += d1[(size_t)s * 64 + o];
    } else if (p < O_B2) {
        const int o = (p - O_W2) / 64, k = (p - O_W2) % 64;
        for (int s = 0; s < n; ++s) g = fmaf(d2[(size_t)s * 64 + o], acts[(size_t)s * 128 + k], g);
    } else if (p < O_W3) {
        const int o = p - O_B2;
        for (int s = 0; s < n; ++s) g += d2[(size_t)s * 64 + o];
    } else if (p < O_B3) {
        const int o = (p - O_W3) / 64, k = (p - O_W3) % 64;
        for (int s = 0; s < n; ++s)
            g = fmaf(d3[(size_t)s * 2 + o], acts[(size_t)s * 128 + 64 + k], g);
    } else {
        const int o = p - O_B3;
        for (int s = 0; s < n; ++s) g += d3[(size_t)s * 2 + o];
    }
    grad[p] = g;
}

// Adam on the FP32 master weights; refreshes the BF16 copy.
__global__ __launch_bounds__(256) void k_adam(float* w, float* m, float* v, const float* grad,
                                               uint16_t* wbf, float lr, float b1, float b2,
                                               float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * g;
    const float vp = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    const float step = lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p] -= step;
    wbf[p] = f32_to_bf16(w[p]);
}

__global__ void k_to_bf16(const float* w, uint16_t* wbf, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) wbf[p] = f32_to_bf16(w[p]);
}

// ------------------------------------------------------------ launchers ----
int features_slabs(int tile, int sub, int dtype) {
    return ((tile + sub - 1) / sub + slab_rows(dtype) - 1) / slab_rows(dtype);
}

hipError_t launch_tile_features_batch(const Descs& ds, const uint32_t* tile_base, int ntiles,
                                      bool v16, int dtype, int conn, int tile, int sub,
                                      float* part, float* raw, float* feat, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    const int nslab = features_slabs(tile, sub, dtype);
    const int grid = ntiles * nslab;
    if (dtype == 1) {
        if (v16) hipLaunchKernelGGL((k_tile_features_part<uint8_t, 16>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
        else hipLaunchKernelGGL((k_tile_features_part<uint8_t, 4>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
    } else {
        if (v16) hipLaunchKernelGGL((k_tile_features_part<uint16_t, 16>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
        else hipLaunchKernelGGL((k_tile_features_part<uint16_t, 4>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
    }
    hipLaunchKernelGGL(k_features_tiles, (ntiles + 3) / 4, 256, 0, s, part, nslab, ntiles, raw);
    hipLaunchKernelGGL(k_features_finish, tile_base ? ds.n : 1, 256, 0, s, ds, tile_base, ntiles,
                       tile, sub, dtype, conn, raw, feat);
    return hipGetLastError();
}

hipError_t launch_tile_features(const Image& im, int dtype, int conn, int tile, int sub,
                                float* part, float* raw, float* feat, hipStream_t s) {
    const int nt = ((im.W + tile - 1) / tile) * ((im.H + tile - 1) / tile);
    Descs ds{};
    ds.n = 1;
    ds.one.J = im.J;
    ds.one.I = im.I;
    ds.one.pj = im.pj;
    ds.one.pi = im.pi;
    ds.one.W = im.W;
    ds.one.H = im.H;
    const bool v16 = (uintptr_t)im.J % 16 == 0 && (uintptr_t)im.I % 16 == 0 && im.pj % 16 == 0 &&
                     im.pi % 16 == 0;
    return launch_tile_features_batch(ds, nullptr, nt, v16, dtype, conn, tile, sub, part, raw,
                                      feat, s);
}

hipError_t launch_plan_features(const Descs& ds, const uint32_t* tile_base, int ntiles, bool v16,
                                int dtype, int tile, int sub, float* part, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    const int nslab = features_slabs(tile, sub, dtype);
    const int grid = ntiles * nslab;
    if (dtype == 1) {
        if (v16) hipLaunchKernelGGL((k_tile_features_part<uint8_t, 16>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
        else hipLaunchKernelGGL((k_tile_features_part<uint8_t, 4>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
    } else {
        if (v16) hipLaunchKernelGGL((k_tile_features_part<uint16_t, 16>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
        else hipLaunchKernelGGL((k_tile_features_part<uint16_t, 4>), grid, 256, 0, s, ds, tile_base, tile, sub, nslab, part);
    }
    return hipGetLastError();
}

hipError_t launch_plan_predict(const Descs& ds, const uint32_t* tile_base, int ntiles, int dtype,
                               int conn, int tile, int sub, const float* part, float* raw,
                               const void* params, float* out, uint32_t* count, uint32_t* flag,
                               uint32_t seq, uint32_t* order, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (order && (tile_base || ntiles > kOrderMax)) return hipErrorInvalidValue;
    const int grid = tile_base ? ds.n : (ntiles + kPlanTiles - 1) / kPlanTiles;
    // raw: statistics (ntiles x 16) and outputs (ntiles x 2) of units too
    // large for LDS, then the priorities (ntiles) shared by the workgroups
    float* prio = raw + (size_t)ntiles * 18;
    hipLaunchKernelGGL(k_plan_predict, grid, kPlanThreads, 0, s, ds, tile_base, ntiles,
                       features_slabs(tile, sub, dtype), tile, sub, dtype, conn, part, raw, params, out,
                       count, flag, seq, prio, order);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const void* params, bool bf16, const float* x, int n, float* y,
                              float* acts, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int grid = std::min(2048, (n + 3) / 4);  // 4 waves (samples) per workgroup
    if (bf16) hipLaunchKernelGGL(k_mlp_forward<true>, grid, 256, 0, s, params, x, n, y, acts);
    else hipLaunchKernelGGL(k_mlp_forward<false>, grid, 256, 0, s, params, x, n, y, acts);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const void* params, bool bf16, const float* x, const float* acts,
                               const float* y, const float* t, int n, float lw0, float lw1,
                               float* d3, float* d2, float* d1, float* grad, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int grid = std::min(1024, (n + 255) / 256);
    if (bf16)
        hipLaunchKernelGGL(k_mlp_backward<true>, grid, 256, 0, s, params, acts, y, t, n, lw0, lw1,
                           d3, d2, d1);
    else
        hipLaunchKernelGGL(k_mlp_backward<false>, grid, 256, 0, s, params, acts, y, t, n, lw0,
                           lw1, d3, d2, d1);
    hipLaunchKernelGGL(k_mlp_grads, (kMlpParams + 255) / 256, 256, 0, s, x, acts, d3, d2, d1, n,
                       grad);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, float* m, float* v, const float* grad, uint16_t* wbf, float lr,
                       int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, (kMlpParams + 255) / 256, 256, 0, s, w, m, v, grad, wbf, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* w, uint16_t* wbf, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, (n + 255) / 256, 256, 0, s, w, wbf, n);
    return hipGetLastError();
}

}  // namespace iwpp
