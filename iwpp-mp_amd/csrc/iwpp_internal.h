// Internal declarations shared by the HIP kernels (iwpp_kernels.hip) and the
// host driver (iwpp_api.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace iwpp {

constexpr int kMP = 64;          // megapixel edge (queue work unit), pixels
constexpr int kWG = 256;         // threads per workgroup for every kernel here
constexpr int kCtlStride = 32;   // control-block counters sit 128 bytes apart

// Control-block slots (each at index * kCtlStride uint32s).
enum Ctl : int {
    C_HEAD = 0,      // mutex ring: next pop ticket
    C_TAIL,          // mutex ring: next push ticket / seeding count
    C_PENDING,       // mutex: megapixels queued or running
    C_RSIZE,         // bsp: entries in the read queue
    C_WSIZE,         // bsp: entries appended to the write queue
    C_CURSOR,        // bsp: next read-queue index to claim
    C_ROUND,         // bsp: round counter
    C_BAR_COUNT,     // grid barrier arrivals
    C_BAR_GEN,       // grid barrier generation
    C_OVERFLOW,      // a queue append was dropped
    C_TIMEOUT,       // a device-side wait hit its limit
    C_VISITS_LO,     // megapixel visits (low word)
    C_VISITS_HI,     // megapixel visits (high word)
    C_CHANGED,       // sweep passes: some pixel rose
    C_PH_POP,        // phase timers (64-bit, wall-clock ticks summed over waves):
    C_PH_FETCH,      //   waiting for / taking a megapixel, loading it,
    C_PH_CONVERGE,   //   local sweeps, write-back, notification
    C_PH_STORE,
    C_PH_NOTIFY,
    C_NSLOTS
};

// Megapixel state bits (mutex variant).
constexpr uint32_t ST_PENDING = 1u;
constexpr uint32_t ST_RUNNING = 2u;

// Tiles that the sweep passes leave untouched ("walls"): active == nullptr
// means every tile takes part.
struct TileMask {
    const uint8_t* active;
    int tw, th, ntx;
};

struct Image {
    uint8_t* J;         // marker (in/out)
    const uint8_t* I;   // mask
    size_t pj, pi;      // pitches in bytes
    int W, H;
};

struct QueueArgs {
    uint32_t* ctl;      // control block
    uint32_t* st;       // per-megapixel state / dedup flag
    uint32_t* qa;       // queue A (mutex ring slots / bsp read queue on even rounds)
    uint32_t* qb;       // queue B (bsp write queue on even rounds)
    unsigned long long* ring;  // mutex ring: (sequence << 32) | megapixel per slot
    uint32_t cap;       // entries per queue (mutex: power of two)
    int nmx, nmy;       // megapixel grid
    int max_iters;      // local sweep rounds per visit
    int max_rounds;     // bsp round limit
    unsigned long long timeout_ticks;  // wall-clock ticks a wait may last
    int trace;          // accumulate the phase timers
};

enum QueueKind : int { Q_MUTEX = 0, Q_DQ = 1, Q_UQ = 2 };

// ---- launchers (iwpp_kernels.hip); each returns a hipError_t ----
hipError_t launch_clip(const Image& im, int dtype, hipStream_t s);
hipError_t launch_row_pass(const Image& im, int dtype, bool clip, const TileMask& tm,
                           uint32_t* changed, hipStream_t s);
// block_rows <= 0 picks the block height from the image size.
hipError_t launch_col_pass(const uint8_t* src, uint8_t* dst, const uint8_t* mask, size_t ps,
                           size_t pd, size_t pi, int W, int H, int dtype, int conn, int dir,
                           int block_rows, const TileMask& tm, uint32_t* changed,
                           hipStream_t s);
int col_block_width(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
                    size_t pi, int W, int dtype);
hipError_t launch_unstable(const Image& im, int dtype, int conn, uint8_t* flags,
                           uint32_t* count, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_ring_init(unsigned long long* ring, size_t n, hipStream_t s);
hipError_t launch_seed_queue(const uint8_t* flags, int nmp, int kind, const QueueArgs& q,
                             hipStream_t s);
hipError_t launch_prepare_queue(const QueueArgs& q, hipStream_t s);
hipError_t launch_queue_phase(const Image& im, int dtype, int conn, int kind, const QueueArgs& q,
                              int workers_per_cu, hipStream_t s);

}  // namespace iwpp
