/*
 * iwpp.h - C ABI of the MI355X-native IWPP grayscale morphological
 * reconstruction by dilation (library: iwpp-mp_amd/libiwpp.so, gfx950).
 *
 * The reference (oliveira-mathias/IWPP-MP) exposes one synchronous entry
 * point, built three times (one per queue variant):
 *   void MorphologicalReconstruction(unsigned char* marker, unsigned char* mask,
 *                                    size_t pitchMarker, size_t pitchMask, int size);
 *   IWPP MP/include/MorphologicalReconstruction.cuh:45-46 (mutex queue),
 *   IWPP MP DQ/include/MorphologicalReconstruction.cuh:53-54 (double queue),
 *   IWPP MP UQ/include/MorphologicalReconstruction.cuh:59-60 (unique queue).
 * This header keeps that exact function (at the bottom) and adds the general
 * entry point the reference hard-codes at build time: dtype, connectivity,
 * width x height, tile size, queue variant and schedule become arguments.
 *
 * Conventions
 *  - Images are device pointers (hipMalloc / hipMallocPitch / torch) with row
 *    pitches in bytes.  Base pointers must be 4-byte aligned and pitches a
 *    multiple of 4 (cudaMallocPitch/hipMallocPitch give both); 16-byte
 *    alignment of base and pitch selects the wide-vector kernels.
 *  - marker is updated in place to the reconstruction; mask is only read.
 *    marker is clipped to mask first (IWPP MP/src/main.cu:55 does this in the
 *    caller; here it is part of the call and is a no-op on clipped input).
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are ordered on
 *    the stream; entry points that report counters or restart after a queue
 *    overflow synchronise the stream before returning.
 *  - Status codes instead of printf: IWPP_OK or a negative IWPP_ERR_*;
 *    iwpp_last_error() gives a message for the calling thread.
 *  - Workspaces own all scratch memory; one workspace per concurrent stream.
 *    ws = NULL uses a per-device default workspace (not thread-safe).
 */
#ifndef IWPP_H
#define IWPP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { IWPP_U8 = 1, IWPP_U16 = 2 };
enum {
    IWPP_QUEUE_MUTEX = 0, /* one global queue, persistent kernel (IWPP MP/src/globalQueue.cu) */
    IWPP_QUEUE_DQ = 1,    /* read/write queues swapped between rounds (IWPP MP DQ/src/globalQueue.cu) */
    IWPP_QUEUE_UQ = 2,    /* DQ with one entry per tile per round (IWPP MP UQ/src/MorphologicalReconstruction.cu:1152-1293) */
    IWPP_QUEUE_AUTO = 3   /* autotuner choice */
};
enum { IWPP_SCHED_PLAIN = 0, IWPP_SCHED_LEARNED = 1 };
enum {
    IWPP_OK = 0,
    IWPP_ERR_ARG = -1,        /* bad argument (null, size, alignment, enum) */
    IWPP_ERR_HIP = -2,        /* a HIP runtime call failed */
    IWPP_ERR_TIMEOUT = -3,    /* a device-side wait exceeded its time limit */
    IWPP_ERR_UNSUPPORTED = -4,/* valid but unsupported combination */
    IWPP_ERR_NOMEM = -5       /* workspace allocation failed */
};

/* Megapixel (queue work unit) edge in pixels: 64 x 64. */
#define IWPP_MP_SIZE 64

typedef struct iwpp_workspace iwpp_workspace;

/* Counters of one reconstruction (filled when a non-NULL pointer is passed). */
typedef struct {
    long long tile_visits;   /* megapixel visits in the queue phase */
    long long seeded;        /* megapixels put in the queue by the seeding pass(es) */
    int sweep_rounds;        /* raster/anti-raster rounds before the queue phase */
    int queue_rounds;        /* BSP rounds (DQ/UQ) */
    int restarts;            /* queue overflows that re-seeded from the partial marker */
    int queue;               /* queue variant actually used */
    int tile;                /* tile size actually used */
    int sweeps_skipped_tiles;/* tile-rounds skipped by the schedule */
    int schedule;            /* schedule actually used (LEARNED needs a model) */
    float ms_sweep;          /* device time of the sweep phase (HIP events) */
    float ms_seed;           /* device time of the seeding passes */
    float ms_queue;          /* device time of the queue-phase kernels */
    float ms_plan;           /* host time the learned schedule waited for the
                                network's outputs (the tile features, the first
                                row pass and the network run meanwhile) */
    /* With iwpp_options.trace_phases: wave-microseconds spent in the queue
     * phase, summed over all waves (the reference's clock64 split,
     * IWPP MP/src/MorphologicalReconstruction.cu:199-231,534-540). */
    double us_wait;          /* taking a megapixel (incl. idle waiting, barriers) */
    double us_fetch;         /* loading megapixel + halo into LDS */
    double us_converge;      /* local sweeps */
    double us_store;         /* write-back */
    double us_notify;        /* neighbour tests and queue pushes */
    /* Split of ms_sweep (HIP events around each pass).  Split sweeps: the
     * row passes, and the column passes (downward and upward, the learned
     * schedule's network kernel excluded).  Raster sweeps: the downward
     * passes, and the upward passes. */
    float ms_rows;
    float ms_cols;
    /* With iwpp_options.check_queue (the reference's queue self-checks,
     * IWPP MP DQ/src/globalQueue.cu:218-253 and IWPP MP UQ/src/
     * MorphologicalReconstruction.cu:1312-1322, as counters):
     * repeat_visits: mutex -- visits that overlapped another visit of the
     * same megapixel; DQ / UQ -- visits of a megapixel already visited in the
     * same round (UQ: must be 0; DQ allows them).  state_left: megapixels
     * whose state word (mutex: PENDING/RUNNING bits; UQ: queued flag) is not
     * clear when the queue phase ends (must be 0). */
    long long repeat_visits;
    long long state_left;
    int col_block_rows;      /* rows per block of the column passes used */
    int seed_class;          /* seed class the autotuner cache was looked up with
                                (-1: not measured, no cached tuning applies) */
    float ms_seams;          /* raster sweeps: device time of the fix-up passes
                                beside the block boundaries (in ms_sweep, not in
                                ms_rows / ms_cols) */
    int seam_rows;           /* rows of those fix-up windows asked for (each is
                                at most half a block; 0 = none) */
} iwpp_stats;

/* Tuning knobs of iwpp_reconstruct_ex; iwpp_default_options() fills them. */
typedef struct {
    int sweep_rounds;     /* sweep rounds (see split_sweeps) before seeding the
                             queue; -1 = the schedule decides */
    int col_block_rows;   /* rows per block of the column passes (0 = from the
                             image size, default) */
    int max_local_iters;  /* local sweep rounds per megapixel visit before it
                             re-queues itself (default 8) */
    int queue_capacity;   /* entries per queue; 0 = sized from the image.
                             Small values force the overflow/restart path. */
    int workers_per_cu;   /* persistent workgroups per CU; 0 = from occupancy */
    double timeout_s;     /* device-side wait limit (default 60 s) */
    int trace_phases;     /* 1 = fill the us_* phase timers of iwpp_stats */
    int check_queue;      /* 1 = check the queue variant's discipline and fill
                             repeat_visits / state_left of iwpp_stats (costs
                             two atomics per visit and one extra launch) */
    long long learned_min_pixels; /* IWPP_SCHED_LEARNED plans a single image only
                             from this many pixels on (default 2^27; smaller
                             images keep the plain rule: the plan's fixed cost
                             exceeds what it can save there); 0 = always */
    int split_sweeps;     /* 0 (default): a sweep round is two raster passes over
                             column blocks, downward then upward, each row
                             scanned across its strip after its column step;
                             1: a row pass (forward and backward) and two
                             column passes (down, up) */
    int seam_rows;        /* raster sweeps: after each raster pass, a short pass
                             of this many rows beside every block boundary
                             (at most half a block), whose first row reads
                             the other block's final row; 0 = none, negative
                             (default -1) = the library's choice */
} iwpp_options;

void iwpp_default_options(iwpp_options* opt);

/* Workspace sized for images up to max_pixels pixels of up to max_dtype bytes,
 * on the current device.  It grows on demand when a call needs more. */
iwpp_workspace* iwpp_workspace_create(size_t max_pixels, int max_dtype);
void iwpp_workspace_destroy(iwpp_workspace* ws);

/* The general entry point (replaces MorphologicalReconstruction, see top).
 * width, height >= 1; dtype IWPP_U8/U16; connectivity 4 or 8;
 * tile = scheduling tile edge in pixels (a multiple of 64, or 0 = auto);
 * queue IWPP_QUEUE_*; schedule IWPP_SCHED_*. */
int iwpp_reconstruct(void* marker, const void* mask, size_t pitch_marker, size_t pitch_mask,
                     int width, int height, int dtype, int connectivity, int tile, int queue,
                     int schedule, iwpp_workspace* ws, void* stream);

/* Same, with options (NULL = defaults) and counters (NULL = not wanted). */
int iwpp_reconstruct_ex(void* marker, const void* mask, size_t pitch_marker, size_t pitch_mask,
                        int width, int height, int dtype, int connectivity, int tile, int queue,
                        int schedule, const iwpp_options* opt, iwpp_stats* stats,
                        iwpp_workspace* ws, void* stream);

/* One image of a batch: device pointers, pitches in bytes, size. */
typedef struct {
    void* marker;
    const void* mask;
    size_t pitch_marker, pitch_mask;
    int width, height;
} iwpp_image;

/* Batched reconstruction: n images of one dtype and connectivity, each in
 * place, in one set of launches (one persistent queue kernel pulls
 * (image, megapixel) work items for all of them).  No reference
 * counterpart: the reference reconstructs one square image per call
 * (IWPP MP/include/MorphologicalReconstruction.cuh:45-46); this is that call
 * for many images at once. */
int iwpp_reconstruct_batch(const iwpp_image* imgs, int n, int dtype, int connectivity, int tile,
                           int queue, int schedule, const iwpp_options* opt, iwpp_stats* stats,
                           iwpp_workspace* ws, void* stream);

/* A batch reconstruction captured once into a hipGraph and replayed.  The
 * capture records every launch of one iwpp_reconstruct_batch (mutex queue
 * variant only: the DQ/UQ kernels use cooperative launches) on `stream`,
 * which must not be the NULL stream; ws must be explicit and must not be
 * reused by other calls while the graph exists.  A
 * replay reconstructs the markers in place: the caller restores the input
 * markers before each replay.  iwpp_batch_graph_stats synchronises `stream`
 * and returns the counters of the last replay. */
typedef struct iwpp_graph iwpp_graph;
iwpp_graph* iwpp_batch_graph_create(const iwpp_image* imgs, int n, int dtype, int connectivity,
                                    int tile, int queue, int schedule, const iwpp_options* opt,
                                    iwpp_workspace* ws, void* stream);
int iwpp_batch_graph_launch(iwpp_graph* g, void* stream);
int iwpp_batch_graph_stats(iwpp_graph* g, iwpp_stats* stats, void* stream);
void iwpp_batch_graph_destroy(iwpp_graph* g);

/* Autotuner.  Times every (queue variant, sweep rounds 0..4) on these very
 * images, then, for the winner, the column passes' block height (128, 256,
 * 512, 1024 rows against the size rule; the sweep passes' tile), best of
 * `trials` runs each (HIP events); restores the input markers afterwards,
 * and remembers the winner for this workload class (dtype, connectivity,
 * single/batch, log2 pixels, log2 widest side, seed class of the markers:
 * sparse seeds / partial / dense): later calls with queue = IWPP_QUEUE_AUTO,
 * sweep_rounds = -1 or col_block_rows = 0 use it.  No reference counterpart
 * (the reference fixes everything at build time). */
typedef struct {
    int queue;
    int sweep_rounds;
    int tile;
    float ms;            /* best measured time of one reconstruction */
    int col_block_rows;  /* 0 = the size rule's height won */
    int seed_class;      /* 0 sparse seeds (< 1% of pixels), 1 partial, 2 dense */
} iwpp_tuning;
int iwpp_autotune(const iwpp_image* imgs, int n, int dtype, int connectivity, int trials,
                  const iwpp_options* opt, iwpp_tuning* best, iwpp_workspace* ws, void* stream);
void iwpp_autotune_clear(void);

/* Learned schedule (IWPP_SCHED_LEARNED).  No reference counterpart: the
 * reference hard-codes its sweep/queue split.  A small MLP (16 tile features
 * -> 64 -> 64 -> 2, ReLU) predicts for each tile of `tile` x `tile` pixels
 * (0) how many sweep rounds it takes before the queue phase (rounded,
 * clamped to 0..4) and (1) its seeding priority.  Inference runs with BF16
 * weights and FP32 accumulation.  The schedule changes only which tiles the
 * sweeps touch and the order in which megapixels enter the queue, never the
 * result.  Without a model IWPP_SCHED_LEARNED behaves as IWPP_SCHED_PLAIN.
 *
 * Parameter layout (IWPP_MLP_PARAMS floats): W1[64][16] b1[64] W2[64][64]
 * b2[64] W3[2][64] b3[2]. */
#define IWPP_MLP_PARAMS 5378
#define IWPP_MLP_FEATURES 16

/* Installs the model for this process (host FP32 weights, rounded to BF16;
 * n must be IWPP_MLP_PARAMS).  weights = NULL removes it. */
int iwpp_sched_set_model(const float* weights, int n);

/* Tile features of one image (device pointers), as a reconstruction with
 * IWPP_SCHED_LEARNED computes them (there with sub = 8): raw[t][16]
 * statistics over every `sub`-th row and feat[t][16] the normalised features
 * the model reads, t over ceil(h/tile) x ceil(w/tile) tiles in row-major
 * order.  Scratch comes from ws (NULL: the device's default workspace). */
int iwpp_tile_features(const void* marker, const void* mask, size_t pitch_marker,
                       size_t pitch_mask, int width, int height, int dtype, int connectivity,
                       int tile, int sub, float* raw, float* feat, iwpp_workspace* ws,
                       void* stream);

/* MLP on device buffers.  params: IWPP_MLP_PARAMS values, BF16 (uint16) when
 * bf16 != 0 else FP32.  y[n][2] = MLP(x[n][16]); acts (may be NULL) receives
 * the hidden activations h1[64] h2[64] per sample for the backward pass. */
int iwpp_mlp_forward(const void* params, int bf16, const float* x, int n, float* y, float* acts,
                     void* stream);

/* Gradient of L = sum_s (lw0 (y0-t0)^2 + lw1 (y1-t1)^2) / (2n) with respect to
 * the parameters (grad: IWPP_MLP_PARAMS floats), from the forward pass's y
 * and acts.  work: n * 130 floats of scratch. */
int iwpp_mlp_backward(const void* params, int bf16, const float* x, const float* acts,
                      const float* y, const float* target, int n, float lw0, float lw1,
                      float* work, float* grad, void* stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected for step
 * >= 1) on FP32 master weights w with moments m, v; writes the BF16 copy wbf. */
int iwpp_mlp_adam(float* w, float* m, float* v, const float* grad, uint16_t* wbf, float lr,
                  int step, void* stream);

/* marker = min(marker, mask) (IWPP MP/src/MorphologicalReconstruction.cu:827-836). */
int iwpp_clip(void* marker, const void* mask, size_t pitch_marker, size_t pitch_mask, int width,
              int height, int dtype, void* stream);

/* One row pass, forward then backward, in place, over whole rows
 * (the reference's X sweeps, IWPP MP DQ/src/MorphologicalReconstruction.cu:817-912, 949-1040).
 * changed (device int, may be NULL) is OR-ed with 1 when a pixel rose. */
int iwpp_sweep_rows(void* marker, const void* mask, size_t pitch_marker, size_t pitch_mask,
                    int width, int height, int dtype, int* changed, void* stream);

/* One column pass, out of place, in blocks of (64 lanes x 16 or 4 bytes)
 * columns x block_rows rows (block_rows <= 0: chosen from the image size) (the reference's Y sweeps,
 * IWPP MP DQ/src/MorphologicalReconstruction.cu:914-947, 1042-1074).
 * dir +1 = downward, -1 = upward.  Neighbours outside a block are read from src. */
int iwpp_sweep_cols(const void* src, void* dst, const void* mask, size_t pitch_src,
                    size_t pitch_dst, size_t pitch_mask, int width, int height, int dtype,
                    int connectivity, int dir, int block_rows, int* changed, void* stream);

/* One raster pass, out of place, over the blocks of iwpp_sweep_cols: in each
 * block the rows are taken in sweep order (dir +1 downward, -1 upward); a
 * row gets the column step of iwpp_sweep_cols, then a forward and a backward
 * scan across the block's columns that enter from src's pixels just left and
 * right of the block in the same row.  Neighbours outside a block are read
 * from src.  A reconstruction's sweep rounds are two such passes in place
 * (down, up).  No reference counterpart as one pass: it fuses the
 * reference's X and Y sweeps (IWPP MP DQ/src/MorphologicalReconstruction.cu:
 * 817-1074) per block, as Vincent's raster scan does. */
int iwpp_sweep_raster(const void* src, void* dst, const void* mask, size_t pitch_src,
                      size_t pitch_dst, size_t pitch_mask, int width, int height, int dtype,
                      int connectivity, int dir, int block_rows, int* changed, void* stream);

/* Column block width in pixels used by iwpp_sweep_cols for these pointers
 * and this width (16-byte lanes for aligned rows of at least 16 KiB, else
 * 4-byte lanes). */
int iwpp_sweep_cols_block_width(const void* src, const void* dst, const void* mask,
                                size_t pitch_src, size_t pitch_dst, size_t pitch_mask,
                                int width, int dtype);

/* Seeding test: flags[b] (device uint8, ceil(h/64)*ceil(w/64) entries) = 1 when
 * megapixel b holds a pixel that a neighbour can still raise.  Writes the
 * number of such megapixels to *count (device int, may be NULL). */
int iwpp_unstable_megapixels(const void* marker, const void* mask, size_t pitch_marker,
                             size_t pitch_mask, int width, int height, int dtype,
                             int connectivity, uint8_t* flags, int* count, void* stream);

/* One sweep round (two raster passes, down and up, each followed by its
 * fix-up pass beside the block boundaries; with split_sweeps a row pass and
 * two column passes; all in place, clipping first) over n images of
 * one dtype, whose upward pass sets the megapixel seeding flags as a
 * reconstruction's last sweep round does
 * before its queue phase (the test of pairs inside each column block and
 * strip during the pass, then the test across strip edges and block
 * boundaries).  flags (device uint8): one per megapixel, the images' grids
 * (ceil(h/64) x ceil(w/64), row-major) one after another in batch order;
 * 1 = a pixel of that megapixel can still be raised by a neighbour (with
 * fix-up passes, also some megapixels that the upward pass flagged and its
 * fix-up made stable: flags are only ever added).
 * block_rows: column block height (0 = as a reconstruction chooses it,
 * otherwise a multiple of 32).  opt (NULL = defaults): its split_sweeps and
 * seam_rows apply.  No reference counterpart: the reference
 * seeds every megapixel (IWPP MP/src/globalQueue.cu:83-97); this exposes the
 * pass that replaces it, so that its flags can be checked on their own. */
int iwpp_sweep_round_seed(const iwpp_image* imgs, int n, int dtype, int connectivity,
                          int block_rows, const iwpp_options* opt, uint8_t* flags,
                          iwpp_workspace* ws, void* stream);

/* Allocation-free form of iwpp_reconstruct_ex.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  One image, plain schedule, mutex queue variant, default tile;
 * opt (NULL = defaults) must leave queue_capacity at 0.  The image may hold
 * at most IWPP_NOALLOC_MAX_MP megapixels of 64 x 64 pixels (one 32768 x 32768
 * image); larger ones are rejected with IWPP_ERR_ARG before anything runs.
 * Its scratch (control block, megapixel states and flags, queue ring:
 * 5.25 MiB) is a static device array that the library holds on each GPU it
 * runs on until the process ends; the call allocates and frees nothing.
 * Every call on one GPU shares that array, which keeps the last call's
 * contents (each call clears what must start at zero): calls that overlap on
 * one GPU (several host threads or streams) race, and the caller must
 * serialise them.  stream NULL = the NULL stream.  Synchronises the stream
 * before returning (to read the queue phase's status). */
#define IWPP_NOALLOC_MAX_MP 262144
int iwpp_reconstruct_noalloc(void* marker, const void* mask, size_t pitch_marker,
                             size_t pitch_mask, int width, int height, int dtype,
                             int connectivity, const iwpp_options* opt, void* stream);

/* Message for the last error on the calling thread ("" if none). */
const char* iwpp_last_error(void);

/* Legacy drop-in with the reference's exact signature and behaviour:
 * uint8, square size x size, 4-connectivity, synchronous, default stream.
 * Queue variant from env IWPP_QUEUE = mutex | dq | uq (default: uq). */
void MorphologicalReconstruction(unsigned char* marker, unsigned char* mask, size_t pitchMarker,
                                 size_t pitchMask, int size);

#ifdef __cplusplus
}
#endif
#endif
