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
    C_SEEDED,        // mutex ring: entries put in by the seeding pass (tickets 0..C_SEEDED-1)
    C_SCLAIM,        // mutex ring: next seeded ticket to claim (claimed in batches)
    C_REPEAT,        // queue checks: visits that broke the variant's discipline
    C_STATE_LEFT,    // queue checks: megapixel state words not clear after the phase
    C_START,         // bsp: blocks past the start gate
    C_NOTRES,        // bsp: the grid was not co-resident (relaunch it smaller)
    C_SEEDDONE,      // seeding: workgroups of k_seed_queue finished
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

// One image of a launch.  Every kernel that can run over several images
// enumerates global work items (rows, column-pass waves, seeding waves,
// megapixels) and finds the image through the *_base prefix offsets.
struct ImgDesc {
    uint8_t* J;          // marker (in/out)
    const uint8_t* I;    // mask
    uint8_t* P;          // scratch plane for the column passes (may be null)
    size_t pj, pi, pp;   // pitches in bytes
    int W, H;
    int nmx, nmy;        // megapixel grid
    int R;               // column-pass block rows
    int strips;          // column strips at the launch's lane width
    int mp_base;         // first global megapixel index
    int row_base;        // first global row
    int cw_base;         // first column-pass wave
    int sw_base;         // first seeding wave
    TileMask tm;         // walls for the sweep passes
    int rounds = 0x7FFFFFFF;  // sweep rounds this image takes (a batch under the
                              // learned schedule); later rounds skip it
    long long edge_base = 0;  // first word of the image's strip border columns
                              // (COL_IN_PLACE_SEED): strips * 2 * H words
    int se_base = 0;          // first wave of the seeding edge test
};

// Images of a launch: a device array, or (d == nullptr) the single image `one`
// passed by value.  img_of_mp maps a global megapixel to its image (device,
// may be null when n == 1).
struct Descs {
    const ImgDesc* d;
    const uint32_t* img_of_mp;
    // work item -> image tables (several images only): global row, column
    // block wave, seeding wave
    const uint32_t* img_of_row;
    const uint32_t* img_of_cw;
    const uint32_t* img_of_sw;
    int n;
    int total;           // work items of the launch (kernel specific)
    int round = 0;       // sweep round of the launch (0-based)
    // Raster fix-up launches (COL_RASTER, COL_RASTER_SEED): > 0 = each wave
    // passes only over the `seam` rows after (downward) or before (upward)
    // the boundary below its block (at most R / 2 rows); 0 = whole blocks.
    int seam = 0;
    ImgDesc one;
};

struct QueueArgs {
    uint32_t* ctl;      // control block
    uint32_t* st;       // per-megapixel state / dedup flag
    uint32_t* qa;       // queue A (mutex ring slots / bsp read queue on even rounds)
    uint32_t* qb;       // queue B (bsp write queue on even rounds)
    unsigned long long* ring;  // mutex ring: (sequence << 32) | megapixel per slot
    uint32_t cap;       // entries per queue (mutex: power of two)
    int nmp;            // megapixels of all images
    int max_iters;      // local sweep rounds per visit
    int max_rounds;     // bsp round limit
    int grid_div;       // bsp: the launch uses 1/grid_div of the co-resident grid
    unsigned long long start_ticks;  // bsp: wall-clock ticks the start gate waits
    unsigned long long timeout_ticks;  // wall-clock ticks a wait may last
    int trace;          // accumulate the phase timers
    // Queue-discipline checks (nullptr: off): one word per megapixel.  mutex:
    // visits of a megapixel in progress at once (must be 1); DQ / UQ: the
    // last round that visited it (UQ: at most once per round).  Breaches
    // count in C_REPEAT.
    uint32_t* chk;
    // Seeding order by tiles (nullptr = megapixel index order): rank r takes
    // tile order_tiles[r]; its megapixels occupy queue positions
    // order_start[r] .. order_start[r+1]-1, row by row.  One image only.
    const uint32_t* order_tiles;
    const uint32_t* order_start;
    int order_nt, order_ntx, order_mpt, order_nmx, order_nmy;
};

enum QueueKind : int { Q_MUTEX = 0, Q_DQ = 1, Q_UQ = 2 };

// ---- launchers (iwpp_kernels.hip); each returns a hipError_t ----
hipError_t launch_clip(const Image& im, int dtype, hipStream_t s);
hipError_t launch_row_pass(const Image& im, int dtype, bool clip, const TileMask& tm,
                           uint32_t* changed, hipStream_t s);
// Row pass over the rows of every image (ds.total = all rows): one wave per
// row when wave_rowbytes > 0 (every image 16-byte aligned, rows of at most
// wave_rowbytes <= 8 KiB), else streamed by one workgroup per row with 16-byte
// (wide) or 4-byte lanes.
hipError_t launch_row_pass_batch(const Descs& ds, int dtype, bool clip, bool wide,
                                 int wave_rowbytes, uint32_t* changed, hipStream_t s);
// Column pass over every image's column blocks (ds.total = all waves).
// mode: COL_J_TO_P read J write P, COL_P_TO_J read P write J, COL_IN_PLACE on
// J; dir +1 down, -1 up; wide = 16-byte lanes (every image 16-byte aligned),
// else 4-byte lanes.
// COL_IN_PLACE_SEED: in place, and sets the megapixel seeding flags (zeroed
// before) as k_unstable_scan would after the pass; without tile schedules,
// the last pass of every image of the launch.  It also writes the strips'
// border columns as they leave the pass: edges[(strip * 2 + side) * H + y] =
// marker | mask << 16 of the strip's first (side 0) or last (side 1) pixel of
// row y (strips * 2 * H words, from edges + edge_base for each image).
// Raster modes (COL_RASTER_*): the same blocks, and after each row's column
// step the row is scanned forward and backward across its strip (the pixels
// just left and right of the strip, in the same row, enter the scans as lower
// bounds), so that one pass propagates down (or up) and sideways:
// COL_RASTER_J_TO_P out of place (every pixel of P written; pixels outside
// the block or strip read from J), COL_RASTER in place, COL_RASTER_CLIP in
// place clipping the marker to the mask first (pixels read from outside the
// block or strip are clipped as they are read), COL_RASTER_SEED in place with
// the seeding flags and border columns of COL_IN_PLACE_SEED.  With ds.seam >
// 0, COL_RASTER and COL_RASTER_SEED make a fix-up pass over the rows beside
// each block boundary, whose first row then reads the other block's final
// row; the seeding form also tests the row beyond its window's far end
// against the window's last row (it may add flags, never clears one).
enum ColMode : int {
    COL_J_TO_P = 0,
    COL_P_TO_J = 1,
    COL_IN_PLACE = 2,
    COL_IN_PLACE_SEED = 3,
    COL_RASTER_J_TO_P = 4,
    COL_RASTER = 5,
    COL_RASTER_CLIP = 6,
    COL_RASTER_SEED = 7
};
// The seeding test across the strip edges (from `edges`) and block
// boundaries of a COL_IN_PLACE_SEED pass, completing its flags (ds.total =
// all edge-test waves, seed_edge_waves() per image from se_base).
hipError_t launch_seed_edges(const Descs& ds, int dtype, int conn, bool wide,
                             const uint32_t* edges, uint8_t* flags, hipStream_t s);
// Edge-test waves of one image (strips and R set).
long seed_edge_waves(const ImgDesc& d);
hipError_t launch_col_pass_batch(const Descs& ds, int dtype, int conn, int dir, int mode,
                                 bool wide, uint32_t* changed, hipStream_t s,
                                 uint8_t* flags = nullptr, uint32_t* edges = nullptr);
bool col_wide(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
              size_t pi, int W, int dtype);
int col_block_width(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
                    size_t pi, int W, int dtype);
int col_auto_rows(int W, int H, int dtype, bool wide);
int col_strips(int W, int dtype, bool wide);
// Seeding test over every image (ds.total = all seeding waves: strips x nmy).
hipError_t launch_unstable_batch(const Descs& ds, int dtype, int conn, bool wide,
                                 uint8_t* flags, uint32_t* count, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
// Work item -> image tables of a batch from its n device descriptors: image
// i's megapixels, rows, column-block waves (strips x blocks of R rows) and
// seeding waves (strips x megapixel rows), from its *_base on, get i.
hipError_t launch_image_tables(const ImgDesc* d, int n, uint32_t* img_of_mp, uint32_t* img_of_row,
                               uint32_t* img_of_cw, uint32_t* img_of_sw, hipStream_t s);
// Compacts the flags into the queue and sets its starting counters (after
// launch_queue_reset).
hipError_t launch_seed_queue(const uint8_t* flags, int nmp, int kind, const QueueArgs& q,
                             hipStream_t s);
// Clears ctl[0..nctl), st[0..nmp) and chk (when not null), and sets ring slot
// i to expect ticket i for i < cap (when ring is not null): one launch.
hipError_t launch_queue_reset(uint32_t* ctl, size_t nctl, uint32_t* st, size_t nmp, uint32_t* chk,
                              unsigned long long* ring, size_t cap, hipStream_t s);
// Autotuner key: counts[0] += pixels with min(marker, mask) > 0, counts[1] +=
// pixels sampled, over every 64th row of every image (counts zeroed before;
// max_h: the tallest image's height).
hipError_t launch_marker_class(const Descs& ds, int dtype, unsigned long long* counts, int max_h,
                               hipStream_t s);
// Counts the nonzero words of q.st[0..nmp) into C_STATE_LEFT (queue checks).
hipError_t launch_state_left(const QueueArgs& q, hipStream_t s);
hipError_t launch_queue_phase(const Descs& ds, int dtype, int conn, int kind,
                              const QueueArgs& q, int workers_per_cu, hipStream_t s);

}  // namespace iwpp
