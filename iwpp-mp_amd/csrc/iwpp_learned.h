// Learned schedule: features and MLP launchers (iwpp_learned.hip).
#pragma once

#include "iwpp_internal.h"

namespace iwpp {

constexpr int kMlpParams = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2;  // 5378
constexpr int kMlpIn = 16, kMlpOut = 2;
constexpr int kFeatureSub = 64;  // row subsampling of the tile features in a reconstruction
constexpr int kOrderMax = 512;  // most tiles whose seeding order the network kernel computes

// BF16 round to nearest even from the IEEE-754 binary32 layout (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu))  // NaN
        return (uint16_t)((u >> 16) | 0x40u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// raw: ntiles x 16 statistics, feat: ntiles x 16 normalised features
// (tiles of `tile` pixels, raster order; every `sub`-th row sampled).
// Slabs per tile of the feature kernel (partials: tiles * slabs * 16 floats).
int features_slabs(int tile, int sub, int dtype);
hipError_t launch_tile_features(const Image& im, int dtype, int conn, int tile, int sub,
                                float* part, float* raw, float* feat, hipStream_t s);
// The same over several images (ds: device descriptor table); image i owns
// the ntiles-wide numbering's tiles tile_base[i] .. tile_base[i+1]-1 (device).
hipError_t launch_tile_features_batch(const Descs& ds, const uint32_t* tile_base, int ntiles,
                                      bool v16, int dtype, int conn, int tile, int sub,
                                      float* part, float* raw, float* feat, hipStream_t s);
// The learned plan's two launches: the feature pass (slab partials of the
// input marker), then the network, which reads only the partials and may run
// after the marker has changed.  `out` (two floats per tile) and `flag` are
// device-visible pinned host memory; `flag` gets `seq` once every output is
// there; `count` is a zeroed device word.  `order` (device, 2 ntiles + 1
// words, one image of at most kOrderMax tiles, else null): the seeding order
// by decreasing predicted priority and each tile's first queue position.
hipError_t launch_plan_features(const Descs& ds, const uint32_t* tile_base, int ntiles, bool v16,
                                int dtype, int tile, int sub, float* part, hipStream_t s);
hipError_t launch_plan_predict(const Descs& ds, const uint32_t* tile_base, int ntiles, int dtype,
                               int conn, int tile, int sub, const float* part, float* raw,
                               const void* params, float* out, uint32_t* count, uint32_t* flag,
                               uint32_t seq, uint32_t* order, hipStream_t s);
// y: n x 2; acts (optional): n x 128 hidden activations for training.
hipError_t launch_mlp_forward(const void* params, bool bf16, const float* x, int n, float* y,
                              float* acts, hipStream_t s);
// Weighted squared loss gradients: grad (kMlpParams) from y - t; d3/d2/d1
// are n x 2 / n x 64 / n x 64 scratch.
hipError_t launch_mlp_backward(const void* params, bool bf16, const float* x, const float* acts,
                               const float* y, const float* t, int n, float lw0, float lw1,
                               float* d3, float* d2, float* d1, float* grad, hipStream_t s);
hipError_t launch_adam(float* w, float* m, float* v, const float* grad, uint16_t* wbf, float lr,
                       int step, hipStream_t s);
hipError_t launch_to_bf16(const float* w, uint16_t* wbf, int n, hipStream_t s);

}  // namespace iwpp
