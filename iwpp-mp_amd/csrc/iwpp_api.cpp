// Host driver and C ABI (include/iwpp.h).  Orchestrates, on one HIP stream:
//   clip (folded into the first row pass) -> sweep rounds (row pass, column
//   pass down, column pass up, all in place) -> seeding
//   (unstable megapixels compacted into the queue) -> queue phase (mutex /
//   DQ / UQ) -> overflow restart from the partially reconstructed marker.
// The restart mirrors the reference's do { ... } while (memLeak) loop
// (IWPP MP/src/MorphologicalReconstruction.cu:785-806,
//  IWPP MP DQ/src/MorphologicalReconstruction.cu:1242-1323): the marker only
// rises and never passes the reconstruction, so any partial marker is a valid
// restart point.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "iwpp.h"
#include "iwpp_internal.h"
#include "iwpp_learned.h"
#include "iwpp_sched.h"

using namespace iwpp;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(IWPP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
size_t next_pow2(size_t v) {
    size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Set while an allocation-free call runs on this thread: grow() then refuses
// to allocate (its buffers are carved out of the static scratch beforehand).
thread_local bool t_no_alloc = false;

template <class P>
hipError_t grow(P*& p, size_t& have, size_t need_elems) {
    if (p && have >= need_elems) return hipSuccess;
    if (t_no_alloc) return hipErrorOutOfMemory;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
        p = nullptr;
        have = 0;
    }
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(need_elems, 1) * sizeof(P));
    if (e == hipSuccess) have = need_elems;
    return e;
}

}  // namespace

struct iwpp_workspace {
    int device = -1;
    uint32_t* st = nullptr;     // per-megapixel state
    size_t st_n = 0;
    uint8_t* flags = nullptr;   // per-megapixel unstable flags
    size_t flags_n = 0;
    uint32_t* edges = nullptr;  // strip border columns of the seeding column pass
    size_t edges_n = 0;
    uint32_t* qa = nullptr;
    size_t qa_n = 0;
    uint32_t* qb = nullptr;
    size_t qb_n = 0;
    uint32_t* chk = nullptr;    // queue-discipline checks (per megapixel)
    size_t chk_n = 0;
    unsigned long long* cls = nullptr;  // seed-class counters (autotuner key)
    size_t cls_n = 0;
    unsigned long long* ring = nullptr;  // mutex ring
    size_t ring_n = 0;
    uint32_t* ctl = nullptr;    // control block
    size_t ctl_n = 0;
    uint32_t* host_ctl = nullptr;  // pinned copy of the control block
    uint8_t* tiles = nullptr;   // per-tile active flags (schedule)
    size_t tiles_n = 0;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // phase timing,
                                                                        // [4]: learned plan ready
    std::vector<hipEvent_t> ev_round;  // per sweep round (kRoundEvents each): raster --
                                       // down pass start, down pass end, up pass start,
                                       // up pass end; split -- row pass start, row pass
                                       // end, column passes start
    ImgDesc* descs = nullptr;   // batch descriptors (device)
    size_t descs_n = 0;
    ImgDesc* host_descs = nullptr;  // pinned staging of their upload
    size_t host_descs_n = 0;
    hipEvent_t descs_read = nullptr;  // recorded after the last upload from it
    uint32_t* img_of_mp = nullptr;  // global megapixel -> image (device)
    size_t img_of_mp_n = 0;
    float* feat_part = nullptr; // learned schedule: per-slab partial statistics
    size_t feat_part_n = 0;
    float* feat_raw = nullptr;  // raw tile statistics
    size_t feat_raw_n = 0;
    uint32_t* order = nullptr;  // seeding order: tile ranks then tile starts
    size_t order_n = 0;
    void* host_plan = nullptr;  // pinned staging for the learned plan (both ways)
    size_t host_plan_n = 0;
    float* host_plan_dev = nullptr;  // the same memory as the device sees it
    uint32_t* plan_flag = nullptr;   // pinned, mapped: the network's "outputs are in" word
    uint32_t* plan_flag_dev = nullptr;
    uint32_t* plan_count = nullptr;  // device word (zero between calls) for that kernel
    uint32_t plan_seq = 0;
    uint16_t* model = nullptr;  // BF16 MLP weights on this device
    int model_version = -1;
    hipStream_t own = nullptr;  // used when the caller passes the NULL stream
};

namespace {
constexpr long long kLearnedMinPixels = 1ll << 27;  // one image: plan from 11585^2 up
// Raster fix-up window (rows beside each block boundary) when the options
// leave it to the library.  Measured on MI355X (one raster round, mutex
// queue): 16384^2 u16 h-dome 0.997 ms without, 0.900 with 16 rows (seeded
// megapixels 12.6K -> 1.4K), 0.948 with 32; 32768^2 u8 seeds 2.154 / 2.144 /
// 2.131 (32K -> 26.5K / 22.5K seeded): each fix-up launch costs about 35 us.
constexpr int kSeamRows = 16;
// Events per timed sweep round (Workspace::ev_round).
constexpr size_t kRoundEvents = 4;
}

extern "C" void iwpp_default_options(iwpp_options* o) {
    if (!o) return;
    o->sweep_rounds = -1;
    o->col_block_rows = 0;
    o->max_local_iters = 8;
    o->queue_capacity = 0;
    o->workers_per_cu = 0;
    o->timeout_s = 60.0;
    o->trace_phases = 0;
    o->check_queue = 0;
    o->learned_min_pixels = kLearnedMinPixels;
    o->split_sweeps = 0;
    o->seam_rows = -1;
}

extern "C" iwpp_workspace* iwpp_workspace_create(size_t max_pixels, int max_dtype) {
    auto* ws = new (std::nothrow) iwpp_workspace();
    if (!ws) return nullptr;
    if (hipGetDevice(&ws->device) != hipSuccess) {
        delete ws;
        return nullptr;
    }
    if (max_pixels && max_dtype >= 1 && max_dtype <= 2) {
        // per-megapixel state, flags and the ring for a single image of
        // max_pixels pixels (edge megapixels of any shape included)
        const size_t nmp = max_pixels / ((size_t)kMP * kMP) + 2 * (size_t)std::sqrt((double)max_pixels) / kMP + 4;
        // strip border columns of the seeding pass for a square image (4-byte
        // lanes, the narrowest strips: 256 bytes)
        const size_t side = (size_t)std::ceil(std::sqrt((double)max_pixels));
        const size_t edge_words = 2 * side * ((side * (size_t)max_dtype + 255) / 256);
        if (grow(ws->st, ws->st_n, nmp) != hipSuccess || grow(ws->flags, ws->flags_n, nmp) != hipSuccess ||
            grow(ws->edges, ws->edges_n, edge_words) != hipSuccess) {
            iwpp_workspace_destroy(ws);
            return nullptr;
        }
    }
    return ws;
}

extern "C" void iwpp_workspace_destroy(iwpp_workspace* ws) {
    if (!ws) return;
    (void)hipFree(ws->st);
    (void)hipFree(ws->flags);
    (void)hipFree(ws->edges);
    (void)hipFree(ws->qa);
    (void)hipFree(ws->qb);
    (void)hipFree(ws->chk);
    (void)hipFree(ws->cls);
    (void)hipFree(ws->ring);
    (void)hipFree(ws->ctl);
    (void)hipFree(ws->tiles);
    (void)hipFree(ws->descs);
    (void)hipFree(ws->img_of_mp);
    (void)hipFree(ws->feat_part);
    (void)hipFree(ws->feat_raw);
    (void)hipFree(ws->order);
    (void)hipFree(ws->model);
    if (ws->host_ctl) (void)hipHostFree(ws->host_ctl);
    if (ws->host_plan) (void)hipHostFree(ws->host_plan);
    if (ws->host_descs) (void)hipHostFree(ws->host_descs);
    if (ws->descs_read) (void)hipEventDestroy(ws->descs_read);
    if (ws->plan_flag) (void)hipHostFree(ws->plan_flag);
    (void)hipFree(ws->plan_count);
    if (ws->own) (void)hipStreamDestroy(ws->own);
    for (hipEvent_t e : ws->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ws->ev_round)
        if (e) (void)hipEventDestroy(e);
    delete ws;
}

namespace {

std::mutex g_default_mu;
std::vector<iwpp_workspace*> g_default_ws;

iwpp_workspace* default_ws() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_default_mu);
    if ((int)g_default_ws.size() <= dev) g_default_ws.resize(dev + 1, nullptr);
    if (!g_default_ws[dev]) g_default_ws[dev] = iwpp_workspace_create(0, 0);
    return g_default_ws[dev];
}

// The stream a call runs on.  For the NULL stream the workspace's own stream
// (created on first use, blocking, so it is ordered after earlier work on the
// NULL stream; the calls that use it synchronise it before returning):
// cooperative launches (DQ/UQ) are refused on the NULL stream by some HIP
// runtimes.
hipError_t call_stream(iwpp_workspace* ws, void* stream, hipStream_t* out) {
    if (stream) {
        *out = (hipStream_t)stream;
        return hipSuccess;
    }
    if (!ws->own) {
        const hipError_t e = hipStreamCreate(&ws->own);
        if (e != hipSuccess) return e;
    }
    *out = ws->own;
    return hipSuccess;
}

// The learned schedule's model: BF16 weights, versioned so that workspaces
// refresh their device copy when it changes.
std::mutex g_model_mu;
std::vector<uint16_t> g_model;
int g_model_version = 0;

int check_image(const void* a, const void* b, size_t pa, size_t pb, int w, int h, int dtype) {
    if (!a || !b) return fail(IWPP_ERR_ARG, "null image pointer");
    if (w <= 0 || h <= 0) return fail(IWPP_ERR_ARG, "bad size %d x %d", w, h);
    if (dtype != IWPP_U8 && dtype != IWPP_U16) return fail(IWPP_ERR_ARG, "bad dtype %d", dtype);
    if ((size_t)w * dtype > (size_t)0x7FFFFFFF || (size_t)h > (size_t)0x7FFFFFFF)
        return fail(IWPP_ERR_ARG, "image too large");
    if (pa < (size_t)w * dtype || pb < (size_t)w * dtype)
        return fail(IWPP_ERR_ARG, "pitch smaller than a row");
    if (pa % 4 || pb % 4) return fail(IWPP_ERR_ARG, "pitch must be a multiple of 4");
    if ((uintptr_t)a % 4 || (uintptr_t)b % 4)
        return fail(IWPP_ERR_ARG, "image pointers must be 4-byte aligned");
    return IWPP_OK;
}

// One reconstruction launch: one image, or several of the same dtype and
// connectivity.
struct Run {
    bool timed;  // record phase times with HIP events
    bool graph;  // graph mode: mutex only, launches without host synchronisation
    bool capturing;  // inside stream capture: no uploads or synchronisation
    int dtype, conn, queue;
    iwpp_options opt;
    iwpp_stats st;
    iwpp_workspace* ws;
    hipStream_t s;
    std::vector<ImgDesc> h;  // host copy of the descriptors
    Descs ds{};              // what the kernels get (totals set per launch)
    bool wide = false;       // 16-byte lanes in the column / seeding passes
    bool row16 = false;      // 16-byte lanes in the batched row pass
    int row_wave = 0;        // batched row pass, one wave per row: the widest row in
                             // bytes (every plane 16-byte aligned), else 0
    int single_row_pass = 0; // one image: use the register row kernels
    int total_rows = 0, total_cw = 0, total_sw = 0, total_mp = 0;
    int total_se = 0;        // seeding edge-test waves
    size_t total_edges = 0;  // words of strip border columns (seeding pass)
    QueueArgs q{};
    size_t cap = 0;
    // learned schedule (single image): per-round tile masks, priority order
    const uint8_t* round_masks = nullptr;
    int sched_tile = 0, sched_ntx = 0, sched_nt = 0;
    const uint32_t* order_tiles = nullptr, *order_start = nullptr;
    int order_nt = 0, order_ntx = 0, order_mpt = 0;
    // learned schedule, between the feature pass and the host's plan
    // seeding flags set by the sweep phase's last pass (one image, no tile
    // schedule): the round whose upward column pass sets them, and whether
    // the flags in the workspace are those of the marker as it now is
    int seed_round = -1;
    bool flags_ready = false;
    int seam = 0;  // rows of the raster fix-up windows (0: none)
    // learned plan still to be completed (the network's outputs read, the
    // rounds fixed): before the first upward column pass, which then sets the
    // seeding flags if the plan keeps to one round
    SchedulePlan* plan_end = nullptr;
    bool plan_predict = false;  // the network is still to be launched (after the first rows)
    bool plan_order = false;    // ... and writes the seeding order to ws->order
    int plan_tiles = 0, plan_tile = 0;
    const uint32_t* plan_tile_base = nullptr;
    uint32_t plan_seq = 0;
    int timed_rounds = 0;  // sweep rounds with per-pass events (ws->ev_round)
    bool scan_seed = false;  // seed from k_unstable_scan only (no strip-edge buffer)
    int seed_class = -1;     // autotuner key of the markers (-1: not measured)
    bool no_tune = false;    // the autotuner's own trials: no cached choices
};

int learned_plan_predict(Run& r);
int learned_plan_end_single(Run& r, SchedulePlan& plan);
int learned_plan_end_batch(Run& r, SchedulePlan& plan);

int ensure_events(iwpp_workspace* ws) {
    for (hipEvent_t& e : ws->ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    return IWPP_OK;
}

int ensure_ctl(iwpp_workspace* ws) {
    HIP_TRY(grow(ws->ctl, ws->ctl_n, (size_t)C_NSLOTS * kCtlStride));
    if (!ws->host_ctl)
        HIP_TRY(hipHostMalloc((void**)&ws->host_ctl, (size_t)C_NSLOTS * kCtlStride * 4, 0));
    return IWPP_OK;
}

unsigned long long timeout_ticks(double seconds) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;  // 100 MHz
    return (unsigned long long)(seconds * khz * 1000.0);
}

// Builds descriptors and the device-side image tables.
int setup_images(Run& r, const iwpp_image* imgs, int n) {
    iwpp_workspace* ws = r.ws;
    r.h.resize(n);
    r.wide = true;
    bool aligned = true;  // every plane 16-byte aligned (base and pitch)
    double useful = 0, lanes16 = 0, chunks16 = 0;
    long strip_rows = 0;
    for (int i = 0; i < n; ++i) {
        ImgDesc& d = r.h[i];
        d = ImgDesc{};
        d.J = (uint8_t*)imgs[i].marker;
        d.I = (const uint8_t*)imgs[i].mask;
        d.P = nullptr;  // the column passes of a reconstruction run in place
        d.pj = imgs[i].pitch_marker;
        d.pi = imgs[i].pitch_mask;
        d.pp = 0;
        d.W = imgs[i].width;
        d.H = imgs[i].height;
        d.nmx = (d.W + kMP - 1) / kMP;
        d.nmy = (d.H + kMP - 1) / kMP;
        d.tm = TileMask{nullptr, kMP, kMP, 1};
        r.wide = r.wide && col_wide(d.J, d.J, d.I, d.pj, d.pj, d.pi, d.W, r.dtype);
        aligned = aligned && (uintptr_t)d.J % 16 == 0 && (uintptr_t)d.I % 16 == 0 &&
                  d.pj % 16 == 0 && d.pi % 16 == 0;
        const double rb = (double)d.W * r.dtype;
        useful += rb * d.H;
        lanes16 += std::ceil(rb / 1024.0) * 1024.0 * d.H;  // 64 lanes x 16 B per column strip
        chunks16 += std::ceil(rb / 4096.0) * 4096.0 * d.H;  // 256 lanes x 16 B per row chunk
    }
    // A batch of narrower images still has plenty of column strips: use the
    // 16-byte lanes whenever they are at least 3/4 busy (and likewise for the
    // row chunks of the batched row pass).
    r.row16 = n > 1 && aligned && useful >= 0.75 * chunks16;
    int widest = 0;
    for (int i = 0; i < n; ++i) widest = std::max(widest, r.h[i].W * r.dtype);
    r.row_wave = aligned ? widest : 0;
    if (n > 1) r.wide = aligned && useful >= 0.75 * lanes16;
    for (int i = 0; i < n; ++i) {
        r.h[i].strips = col_strips(r.h[i].W, r.dtype, r.wide);
        strip_rows += (long)r.h[i].strips * r.h[i].H;
    }
    // One block height for the launch.  In place, a column block carries the
    // rise down (or up) its whole height within one pass, and only across
    // block boundaries by chance, so taller blocks leave the queue less to do;
    // shorter ones give more waves.  Measured on MI355X, best: 512 rows at
    // 32768^2 u8 (2048 waves; queue phase 0.93 -> 0.53 ms against 128 rows),
    // 256 at 16384^2 u16 (2048 waves), 64 at 4096^2 u8 (1024 waves), 32 at
    // 1024^2: about 2048 waves, and at most 64 blocks down a column.  A batch
    // takes at most 256: its images' heights are not multiples of the block
    // height, so its waves' walks differ in length, and the pass lasts as
    // long as its longest walks; more, shorter blocks balance them.  Measured
    // on the 256-image batch (sides 256..4096, 404 Mpixel; 4-byte lanes):
    // 512 rows (3.3K waves) 1.95 ms, 256 (6.7K) 1.88, 128 (13K) 1.99.
    r.seam = r.opt.split_sweeps ? 0 : r.opt.seam_rows < 0 ? kSeamRows : r.opt.seam_rows;
    int R = r.opt.col_block_rows;
    if (R <= 0) {
        int maxh = 0;
        for (int i = 0; i < n; ++i) maxh = std::max(maxh, r.h[i].H);
        long rr = std::max<long>(strip_rows / 2048, maxh / 64);
        rr = (rr + 16) / 32 * 32;
        R = (int)std::min<long>(n > 1 ? 256 : 512, std::max<long>(32, rr));
    }
    long rows = 0, cw = 0, sw = 0, mp = 0, se = 0;
    long long edges = 0;
    for (int i = 0; i < n; ++i) {
        ImgDesc& d = r.h[i];
        d.R = R;
        d.row_base = (int)rows;
        d.cw_base = (int)cw;
        d.sw_base = (int)sw;
        d.mp_base = (int)mp;
        d.se_base = (int)se;
        d.edge_base = edges;
        rows += d.H;
        cw += (long)d.strips * ((d.H + R - 1) / R);
        sw += (long)d.strips * d.nmy;
        mp += (long)d.nmx * d.nmy;
        se += seed_edge_waves(d);
        edges += (long long)d.strips * 2 * d.H;
    }
    if (rows > 0x7FFFFFFF || cw > 0x7FFFFFFF || sw > 0x7FFFFFFF || mp > 0x7FFFFFF0 ||
        se > 0x7FFFFFFF)
        return fail(IWPP_ERR_ARG, "batch too large");
    r.total_se = (int)se;
    r.total_edges = (size_t)edges;
    r.total_rows = (int)rows;
    r.total_cw = (int)cw;
    r.total_sw = (int)sw;
    r.total_mp = (int)mp;
    r.ds = Descs{};
    r.ds.n = n;
    if (n == 1) {
        r.ds.one = r.h[0];
        return IWPP_OK;
    }
    // device tables: descriptors, and work item -> image for megapixels,
    // rows, column-block waves and seeding waves (one allocation), the
    // tables filled on the device from the descriptors' bases and counts
    // (from the host they were megabytes of pageable copies per call)
    const size_t ntab = (size_t)mp + rows + cw + sw;
    if (!r.capturing) {  // a capture reuses the tables its dry run uploaded
        HIP_TRY(grow(ws->descs, ws->descs_n, (size_t)n));
        HIP_TRY(grow(ws->img_of_mp, ws->img_of_mp_n, ntab));
        // The upload goes through the workspace's pinned staging copy, so
        // nothing waits for it here (with the host vector, which dies with
        // this call, the stream had to be synchronised: 25-35 us of idle GPU
        // per call at 256 images).  The staging copy is rewritten only after
        // the previous upload from it has read it.
        if (ws->descs_read) HIP_TRY(hipEventSynchronize(ws->descs_read));
        if (ws->host_descs_n < (size_t)n) {
            if (ws->host_descs) HIP_TRY(hipHostFree(ws->host_descs));
            ws->host_descs = nullptr;
            ws->host_descs_n = 0;
            HIP_TRY(hipHostMalloc((void**)&ws->host_descs, sizeof(ImgDesc) * n, 0));
            ws->host_descs_n = (size_t)n;
        }
        if (!ws->descs_read)
            HIP_TRY(hipEventCreateWithFlags(&ws->descs_read, hipEventDisableTiming));
        std::memcpy(ws->host_descs, r.h.data(), sizeof(ImgDesc) * n);
        HIP_TRY(hipMemcpyAsync(ws->descs, ws->host_descs, sizeof(ImgDesc) * n,
                               hipMemcpyHostToDevice, r.s));
        HIP_TRY(hipEventRecord(ws->descs_read, r.s));
        HIP_TRY(launch_image_tables(ws->descs, n, ws->img_of_mp, ws->img_of_mp + mp,
                                    ws->img_of_mp + mp + rows, ws->img_of_mp + mp + rows + cw,
                                    r.s));
    }
    r.ds.d = ws->descs;
    r.ds.img_of_mp = ws->img_of_mp;
    r.ds.img_of_row = ws->img_of_mp + mp;
    r.ds.img_of_cw = ws->img_of_mp + mp + rows;
    r.ds.img_of_sw = ws->img_of_mp + mp + rows + cw;
    return IWPP_OK;
}

Descs with_total(Descs d, int total, int round = 0) {
    d.total = total;
    d.round = round;
    return d;
}

// The learned plan's end on the host (the network's outputs), while the
// downward pass of the first round runs.
int plan_end_now(Run& r) {
    if (!r.plan_end) return IWPP_OK;
    SchedulePlan& plan = *r.plan_end;
    r.plan_end = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = r.ds.n == 1 ? learned_plan_end_single(r, plan) : learned_plan_end_batch(r, plan);
    if (rc) return rc;
    r.st.ms_plan = (float)std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t0).count();
    r.seed_round = plan.sweep_rounds - 1;
    return IWPP_OK;
}

// Whether round k's last pass sets the seeding flags (k_unstable_scan is
// then skipped): the last round of every image, no tile schedule, a first
// queue attempt.  Clears the flags and sizes the border-column buffer.
int seeding_pass(Run& r, int k, bool* seed) {
    bool s = !r.scan_seed && k == r.seed_round && !r.round_masks;
    for (int i = 0; s && i < r.ds.n; ++i) s = r.h[i].rounds > k;
    *seed = s;
    if (s) {
        const size_t words = ((size_t)r.total_mp + 3) / 4;
        HIP_TRY(grow(r.ws->flags, r.ws->flags_n, words * 4));
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(r.ws->flags), 0u, words, r.s));
        HIP_TRY(grow(r.ws->edges, r.ws->edges_n, std::max<size_t>(1, r.total_edges)));
    }
    return IWPP_OK;
}

// One raster round: a downward raster pass (clipping in round 0) and an
// upward one (setting the seeding flags when it is the last pass).  The
// learned plan's network runs before the downward pass (it reads only the
// tile features), so that the host completes the plan while that pass runs
// and knows before the upward pass whether it is the last.
int raster_round(Run& r, int k, hipEvent_t* ev) {
    const int rp = learned_plan_predict(r);
    if (rp) return rp;
    const Descs ds = with_total(r.ds, r.total_cw, k);
    Descs fix = ds;
    fix.seam = r.seam;
    if (ev) HIP_TRY(hipEventRecord(ev[0], r.s));
    HIP_TRY(launch_col_pass_batch(ds, r.dtype, r.conn, +1, k == 0 ? COL_RASTER_CLIP : COL_RASTER,
                                  r.wide, nullptr, r.s));
    if (ev) HIP_TRY(hipEventRecord(ev[1], r.s));
    if (r.seam > 0)
        HIP_TRY(launch_col_pass_batch(fix, r.dtype, r.conn, +1, COL_RASTER, r.wide, nullptr, r.s));
    if (ev) HIP_TRY(hipEventRecord(ev[2], r.s));
    int rc = plan_end_now(r);
    if (rc) return rc;
    bool seed = false;
    rc = seeding_pass(r, k, &seed);
    if (rc) return rc;
    const int mode = seed ? COL_RASTER_SEED : COL_RASTER;
    uint8_t* flags = seed ? r.ws->flags : nullptr;
    uint32_t* edges = seed ? r.ws->edges : nullptr;
    HIP_TRY(launch_col_pass_batch(ds, r.dtype, r.conn, -1, mode, r.wide, nullptr, r.s, flags, edges));
    if (ev) HIP_TRY(hipEventRecord(ev[3], r.s));
    if (r.seam > 0)
        HIP_TRY(launch_col_pass_batch(fix, r.dtype, r.conn, -1, mode, r.wide, nullptr, r.s, flags,
                                      edges));
    r.flags_ready = seed;
    return IWPP_OK;
}

// Sweep phase: rounds k0..k1-1, each either two raster passes (down, up; the
// default) or a row pass + column pass down + column pass up
// (opt.split_sweeps); round 0 clips the marker to the mask in its first pass.
int sweep_phase(Run& r, int k0, int k1) {
    if (k1 <= 0) {  // clip only
        if (r.ds.n == 1) {
            const ImgDesc& d = r.h[0];
            HIP_TRY(launch_clip(Image{d.J, d.I, d.pj, d.pi, d.W, d.H}, r.dtype, r.s));
        } else {  // a round index past every image's rounds: rows are only clipped
            HIP_TRY(launch_row_pass_batch(with_total(r.ds, r.total_rows, 0x7FFFFFFF), r.dtype,
                                          true, r.row16, r.row_wave, nullptr, r.s));
        }
        return IWPP_OK;
    }
    for (int k = k0; k < k1; ++k) {
        if (r.round_masks) {
            const TileMask tm{r.round_masks + (size_t)k * r.sched_nt, r.sched_tile, r.sched_tile,
                              r.sched_ntx};
            r.h[0].tm = tm;
            r.ds.one.tm = tm;
        }
        hipEvent_t* ev = nullptr;  // this round's pass events (timed runs)
        if (r.timed) {
            std::vector<hipEvent_t>& er = r.ws->ev_round;
            while (er.size() < kRoundEvents * (size_t)(k + 1)) {
                hipEvent_t e = nullptr;
                HIP_TRY(hipEventCreate(&e));
                er.push_back(e);
            }
            ev = er.data() + kRoundEvents * k;
            r.timed_rounds = k + 1;
        }
        if (!r.opt.split_sweeps) {
            const int rs = raster_round(r, k, ev);
            if (rs) return rs;
            continue;
        }
        if (ev) HIP_TRY(hipEventRecord(ev[0], r.s));
        if (r.ds.n == 1) {
            const ImgDesc& d = r.h[0];
            HIP_TRY(launch_row_pass(Image{d.J, d.I, d.pj, d.pi, d.W, d.H}, r.dtype, k == 0, d.tm,
                                    nullptr, r.s));
        } else {
            HIP_TRY(launch_row_pass_batch(with_total(r.ds, r.total_rows, k), r.dtype, k == 0, r.row16,
                                          r.row_wave, nullptr, r.s));
        }
        if (ev) HIP_TRY(hipEventRecord(ev[1], r.s));
        // learned plan: the network goes between the first row pass and the
        // column passes, so the host plans while those run
        const int rp = learned_plan_predict(r);
        if (rp) return rp;
        if (ev) HIP_TRY(hipEventRecord(ev[2], r.s));
        HIP_TRY(launch_col_pass_batch(with_total(r.ds, r.total_cw, k), r.dtype, r.conn, +1,
                                      COL_IN_PLACE, r.wide, nullptr, r.s));
        // the host plans while the downward pass runs
        int rc = plan_end_now(r);
        if (rc) return rc;
        bool seed = false;
        rc = seeding_pass(r, k, &seed);
        if (rc) return rc;
        HIP_TRY(launch_col_pass_batch(with_total(r.ds, r.total_cw, k), r.dtype, r.conn, -1,
                                      seed ? COL_IN_PLACE_SEED : COL_IN_PLACE, r.wide, nullptr,
                                      r.s, seed ? r.ws->flags : nullptr,
                                      seed ? r.ws->edges : nullptr));
        r.flags_ready = seed;
    }
    r.st.sweep_rounds += std::max(0, k1 - k0);
    return IWPP_OK;
}

int setup_queue(Run& r) {
    iwpp_workspace* ws = r.ws;
    const int kind = r.queue;
    const size_t nmp = (size_t)r.total_mp;
    HIP_TRY(grow(ws->st, ws->st_n, nmp));
    HIP_TRY(grow(ws->flags, ws->flags_n, (nmp + 3) / 4 * 4));
    // (sized here too, so that a graph's dry run allocates it before capture)
    if (!r.scan_seed) HIP_TRY(grow(ws->edges, ws->edges_n, std::max<size_t>(1, r.total_edges)));
    size_t cap;
    if (kind == Q_MUTEX) {
        cap = next_pow2(nmp + 16384);
        HIP_TRY(grow(ws->ring, ws->ring_n, cap));
    } else {
        if (r.opt.queue_capacity > 0) cap = (size_t)r.opt.queue_capacity;
        else cap = kind == Q_DQ ? 8 * nmp + 1024 : nmp + 64;
        HIP_TRY(grow(ws->qa, ws->qa_n, cap));
        HIP_TRY(grow(ws->qb, ws->qb_n, cap));
    }
    if (cap > 0xFFFFFFF0ull) return fail(IWPP_ERR_ARG, "queue capacity too large");
    int rc = ensure_ctl(ws);
    if (rc) return rc;
    QueueArgs& q = r.q;
    q = QueueArgs{};
    q.ctl = ws->ctl;
    q.st = ws->st;
    q.qa = ws->qa;
    q.qb = ws->qb;
    q.ring = ws->ring;
    q.cap = (uint32_t)cap;
    q.nmp = r.total_mp;
    q.max_iters = std::max(1, r.opt.max_local_iters);
    q.max_rounds = 1 << 30;
    q.grid_div = 1;
    q.start_ticks = timeout_ticks(0.02);  // 20 ms for the grid to start together
    q.timeout_ticks = timeout_ticks(r.opt.timeout_s > 0 ? r.opt.timeout_s : 60.0);
    q.trace = r.opt.trace_phases ? 1 : 0;
    if (r.opt.check_queue) {
        HIP_TRY(grow(ws->chk, ws->chk_n, nmp));
        q.chk = ws->chk;
    }
    q.order_tiles = r.order_tiles;
    q.order_start = r.order_start;
    q.order_nt = r.order_nt;
    q.order_ntx = r.order_ntx;
    q.order_mpt = r.order_mpt;
    if (r.order_tiles) {
        q.order_nmx = r.h[0].nmx;
        q.order_nmy = r.h[0].nmy;
    }
    r.cap = cap;
    return IWPP_OK;
}

// Seeding + queue kernel launches (no synchronisation).
int queue_launches(Run& r) {
    iwpp_workspace* ws = r.ws;
    const int kind = r.queue;
    if (r.timed) HIP_TRY(hipEventRecord(ws->ev[1], r.s));
    HIP_TRY(launch_queue_reset(ws->ctl, (size_t)C_NSLOTS * kCtlStride, ws->st, (size_t)r.total_mp,
                               r.q.chk, kind == Q_MUTEX ? ws->ring : nullptr, r.cap, r.s));
    if (!r.flags_ready)  // (after a restart the marker has changed: scan)
        HIP_TRY(launch_unstable_batch(with_total(r.ds, r.total_sw), r.dtype, r.conn, r.wide,
                                      ws->flags, nullptr, r.s));
    else  // the last column pass set the flags inside its blocks: the edges
        HIP_TRY(launch_seed_edges(with_total(r.ds, r.total_se), r.dtype, r.conn, r.wide,
                                  ws->edges, ws->flags, r.s));
    r.flags_ready = false;
    HIP_TRY(launch_seed_queue(ws->flags, r.total_mp, kind, r.q, r.s));
    if (r.timed) HIP_TRY(hipEventRecord(ws->ev[2], r.s));
    HIP_TRY(launch_queue_phase(r.ds, r.dtype, r.conn, kind, r.q, r.opt.workers_per_cu, r.s));
    if (r.timed) HIP_TRY(hipEventRecord(ws->ev[3], r.s));
    if (r.q.chk && kind != Q_DQ) HIP_TRY(launch_state_left(r.q, r.s));
    return IWPP_OK;
}

// Reads the control block (after the stream reached it) into the counters.
// Returns 1 when the queue overflowed (restart needed).
int read_counters(Run& r) {
    const uint32_t* h = r.ws->host_ctl;
    const double us_per_tick = 1e6 / ((double)timeout_ticks(1.0));
    r.st.seeded += h[C_SEEDED * kCtlStride] + 0ll;  // (C_TAIL also counts the mutex ring's pushes)
    r.st.tile_visits += (long long)(*(const unsigned long long*)(h + C_VISITS_LO * kCtlStride));
    r.st.queue_rounds += (int)h[C_ROUND * kCtlStride];
    r.st.repeat_visits += h[C_REPEAT * kCtlStride];
    r.st.state_left += h[C_STATE_LEFT * kCtlStride];
    if (r.q.trace) {
        auto ph = [&](int slot) {
            return (double)(*(const unsigned long long*)(h + slot * kCtlStride)) * us_per_tick;
        };
        r.st.us_wait += ph(C_PH_POP);
        r.st.us_fetch += ph(C_PH_FETCH);
        r.st.us_converge += ph(C_PH_CONVERGE);
        r.st.us_store += ph(C_PH_STORE);
        r.st.us_notify += ph(C_PH_NOTIFY);
    }
    if (h[C_TIMEOUT * kCtlStride]) return fail(IWPP_ERR_TIMEOUT, "device-side queue wait timed out");
    return h[C_OVERFLOW * kCtlStride] ? 1 : 0;
}

// Seeding + queue phase, restarted while a queue overflowed.
int queue_phase(Run& r) {
    iwpp_workspace* ws = r.ws;
    int rc = setup_queue(r);
    if (rc) return rc;
    const int max_restarts = 100000;
    for (int attempt = 0;; ++attempt) {
        rc = queue_launches(r);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(ws->host_ctl, ws->ctl, (size_t)C_NSLOTS * kCtlStride * 4,
                               hipMemcpyDeviceToHost, r.s));
        HIP_TRY(hipStreamSynchronize(r.s));
        if (r.timed) {
            float a = 0, b = 0;
            HIP_TRY(hipEventElapsedTime(&a, ws->ev[1], ws->ev[2]));
            HIP_TRY(hipEventElapsedTime(&b, ws->ev[2], ws->ev[3]));
            r.st.ms_seed += a;
            r.st.ms_queue += b;
        }
        if (ws->host_ctl[C_NOTRES * kCtlStride]) {  // not co-resident: nothing ran
            if (r.q.grid_div >= 64) return fail(IWPP_ERR_HIP, "queue kernel never co-resident");
            r.q.grid_div *= 2;
            --attempt;
            continue;
        }
        rc = read_counters(r);
        if (rc < 0) return rc;
        if (rc == 0) break;
        r.st.restarts++;
        if (attempt >= max_restarts) return fail(IWPP_ERR_TIMEOUT, "too many queue restarts");
    }
    return IWPP_OK;
}

int check_common(int dtype, int conn, int tile, int queue, int schedule) {
    if (dtype != IWPP_U8 && dtype != IWPP_U16) return fail(IWPP_ERR_ARG, "bad dtype %d", dtype);
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (tile < 0 || tile % kMP) return fail(IWPP_ERR_ARG, "tile must be 0 or a multiple of 64");
    if (queue < IWPP_QUEUE_MUTEX || queue > IWPP_QUEUE_AUTO)
        return fail(IWPP_ERR_ARG, "bad queue variant %d", queue);
    if (schedule != IWPP_SCHED_PLAIN && schedule != IWPP_SCHED_LEARNED)
        return fail(IWPP_ERR_ARG, "bad schedule %d", schedule);
    return IWPP_OK;
}

// Learned schedule for one image: tile features -> MLP (BF16) -> per tile
// a predicted number of sweep rounds before the queue phase (0..4) and a
// priority.  Round r of the sweep phase touches the tiles scheduled for more
// than r rounds; the queue is seeded tile by tile in decreasing priority.
// Without a model the plain plan is kept.
// The current model on the workspace's device; *have = false without one.
int ensure_model(iwpp_workspace* ws, bool* have) {
    std::vector<uint16_t> model;
    int version = 0;
    {
        std::lock_guard<std::mutex> lk(g_model_mu);
        *have = !g_model.empty();
        if (!*have) return IWPP_OK;
        if (ws->model_version != g_model_version) model = g_model;
        version = g_model_version;
    }
    if (ws->model_version != version) {
        if (!ws->model) HIP_TRY(hipMalloc((void**)&ws->model, kMlpParams * sizeof(uint16_t)));
        HIP_TRY(hipMemcpy(ws->model, model.data(), kMlpParams * sizeof(uint16_t),
                          hipMemcpyHostToDevice));
        ws->model_version = version;
    }
    return IWPP_OK;
}

// Pinned host staging of at least `bytes`.
int ensure_host_plan(iwpp_workspace* ws, size_t bytes) {
    if (ws->host_plan_n >= bytes) return IWPP_OK;
    if (ws->host_plan) HIP_TRY(hipHostFree(ws->host_plan));
    ws->host_plan = nullptr;
    ws->host_plan_n = 0;
    // mapped and coherent: the network's outputs are written here by the
    // device and read by the host once the stream's event has passed
    HIP_TRY(hipHostMalloc(&ws->host_plan, bytes, hipHostMallocMapped | hipHostMallocCoherent));
    ws->host_plan_n = bytes;
    HIP_TRY(hipHostGetDevicePointer((void**)&ws->host_plan_dev, ws->host_plan, 0));
    return IWPP_OK;
}

int ensure_plan_flag(iwpp_workspace* ws) {
    if (ws->plan_flag) return IWPP_OK;
    HIP_TRY(hipMalloc((void**)&ws->plan_count, 4));
    HIP_TRY(hipMemset(ws->plan_count, 0, 4));
    HIP_TRY(hipHostMalloc((void**)&ws->plan_flag, 64, hipHostMallocMapped | hipHostMallocCoherent));
    *(volatile uint32_t*)ws->plan_flag = 0;
    ws->plan_seq = 0;
    HIP_TRY(hipHostGetDevicePointer((void**)&ws->plan_flag_dev, ws->plan_flag, 0));
    return IWPP_OK;
}

// Launches the network of the learned plan: it reads only the feature
// partials of the input marker, so it goes before the first raster pass (or,
// with split sweeps, between the first row pass and the column passes).
int learned_plan_predict(Run& r) {
    if (!r.plan_predict) return IWPP_OK;
    r.plan_predict = false;
    iwpp_workspace* ws = r.ws;
    Descs one{};
    one.n = 1;
    one.one = r.h[0];
    HIP_TRY(launch_plan_predict(r.h.size() == 1 ? one : r.ds, r.plan_tile_base, r.plan_tiles,
                                r.dtype, r.conn, r.plan_tile, kFeatureSub, ws->feat_part,
                                ws->feat_raw, ws->model, ws->host_plan_dev, ws->plan_count,
                                ws->plan_flag_dev, r.plan_seq,
                                r.plan_order ? ws->order : nullptr, r.s));
    return IWPP_OK;
}

// Waits (polling the pinned flag) until the network's outputs are in host
// memory; a failed stream or the time limit ends the wait with an error.
int learned_plan_wait(Run& r) {
    volatile uint32_t* flag = r.ws->plan_flag;
    const auto t0 = std::chrono::steady_clock::now();
    const double limit = r.opt.timeout_s > 0 ? r.opt.timeout_s : 60.0;
    for (uint32_t k = 1; *flag != r.plan_seq; ++k) {
        if ((k & 1023u) == 0) {
            const hipError_t e = hipStreamQuery(r.s);
            if (e != hipSuccess && e != hipErrorNotReady)
                return fail(IWPP_ERR_HIP, "learned plan: %s", hipGetErrorString(e));
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > limit)
                return fail(IWPP_ERR_TIMEOUT, "learned plan: the network's outputs did not arrive");
            if (e == hipSuccess && *flag != r.plan_seq)
                return fail(IWPP_ERR_HIP, "learned plan: stream idle without the network's outputs");
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return IWPP_OK;
}

int clamp_rounds(float v) {
    return std::isfinite(v) ? (int)std::lround(std::min(4.f, std::max(0.f, v))) : 0;
}

// Learned schedule, in two halves around the first sweep round.  begin
// launches the tile features, the network and the read-back of its
// predictions, and records an event; the first sweep round (needed by every
// schedule that sweeps at all, and harmless otherwise) is launched behind it;
// end waits for the event, so the host plans while that round runs, and
// fixes the remaining rounds.  Every tile of every image is one network
// sample; image i runs R_i = round(area-weighted mean of its tiles'
// predicted rounds) rounds in total.

// Tile counts per image (prefix), into pinned staging after the predictions.
struct LearnedLayout {
    int tile = 0, total = 0;
    std::vector<uint32_t> base;  // n + 1
};

LearnedLayout learned_layout(const Run& r, int tile) {
    LearnedLayout L;
    const int n = (int)r.h.size();
    L.tile = tile;
    L.base.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const ImgDesc& d = r.h[i];
        L.base[i + 1] = L.base[i] + (uint32_t)(((d.W + tile - 1) / tile) * ((d.H + tile - 1) / tile));
    }
    L.total = (int)L.base[n];
    return L;
}

// Pinned staging: predictions (2 floats per tile), then per kind: the batch
// keeps tile bases and the descriptors; one image keeps tile order, starts
// and per-round masks.
size_t learned_staging_bytes(const Run& r, int total) {
    const size_t n = r.h.size();
    return (size_t)total * 2 * sizeof(float) + (n + 1) * 4 + n * sizeof(ImgDesc) +
           (size_t)(2 * total + 1) * 4 + (size_t)4 * total + 256;
}

int learned_plan_begin(Run& r, const SchedulePlan& plan, bool* pending) {
    *pending = false;
    iwpp_workspace* ws = r.ws;
    bool have = false;
    int rc = ensure_model(ws, &have);
    if (rc || !have) return rc;  // no model: the plain heuristic
    rc = ensure_events(ws);
    if (rc) return rc;
    const int n = (int)r.h.size();
    const LearnedLayout L = learned_layout(r, plan.tile);
    rc = ensure_host_plan(ws, learned_staging_bytes(r, L.total));
    if (rc) return rc;
    rc = ensure_plan_flag(ws);
    if (rc) return rc;
    float* pred = (float*)ws->host_plan;
    HIP_TRY(grow(ws->feat_part, ws->feat_part_n,
                 (size_t)L.total * features_slabs(L.tile, kFeatureSub, r.dtype) * 16));
    HIP_TRY(grow(ws->feat_raw, ws->feat_raw_n, (size_t)L.total * 19));  // statistics, outputs, priorities
    if (n == 1) {
        const ImgDesc& d = r.h[0];
        Descs one{};
        one.n = 1;
        one.one = d;
        const bool v16 = (uintptr_t)d.J % 16 == 0 && (uintptr_t)d.I % 16 == 0 && d.pj % 16 == 0 &&
                         d.pi % 16 == 0;
        HIP_TRY(launch_plan_features(one, nullptr, L.total, v16, r.dtype, L.tile, kFeatureSub,
                                     ws->feat_part, r.s));
        r.plan_tile_base = nullptr;
        r.plan_order = false;  // seeding keeps index order (see learned_plan_end_single)
    } else {
        bool v16 = true;
        for (const ImgDesc& d : r.h)
            v16 = v16 && (uintptr_t)d.J % 16 == 0 && (uintptr_t)d.I % 16 == 0 && d.pj % 16 == 0 &&
                  d.pi % 16 == 0;
        uint32_t* h_base = (uint32_t*)(pred + (size_t)L.total * 2);
        std::copy(L.base.begin(), L.base.end(), h_base);
        HIP_TRY(grow(ws->order, ws->order_n, (size_t)n + 1));
        HIP_TRY(hipMemcpyAsync(ws->order, h_base, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, r.s));
        HIP_TRY(launch_plan_features(r.ds, ws->order, L.total, v16, r.dtype, L.tile,
                                     kFeatureSub, ws->feat_part, r.s));
        r.plan_tile_base = ws->order;
    }
    r.plan_tiles = L.total;
    r.plan_tile = L.tile;
    r.plan_seq = ++ws->plan_seq;
    if (r.plan_seq == 0) r.plan_seq = ++ws->plan_seq;  // 0 is the flag's initial value
    r.plan_predict = true;  // launched after the first row pass (sweep_phase)
    *pending = true;
    return IWPP_OK;
}

// Area-weighted mean of the predicted rounds of image i's tiles.
int image_rounds(const ImgDesc& d, const float* pred, uint32_t b0, uint32_t b1, int tile) {
    const int ntx = (d.W + tile - 1) / tile;
    double wsum = 0, wy = 0;
    for (uint32_t g = b0; g < b1; ++g) {
        const int t = (int)(g - b0), tx = t % ntx, ty = t / ntx;
        const double area = (double)(std::min(d.W, (tx + 1) * tile) - tx * tile) *
                            (double)(std::min(d.H, (ty + 1) * tile) - ty * tile);
        const float v = pred[(size_t)g * 2];
        wy += area * (std::isfinite(v) ? std::min(4.f, std::max(0.f, v)) : 0.f);
        wsum += area;
    }
    return clamp_rounds((float)(wy / wsum));
}

// Batch: the sweep kernels skip image i from round R_i on; seeding keeps
// index order.  R_i is capped at the batch's plain (or autotuned) rounds: the
// network's rounds come from images timed one at a time, where the queue
// phase of a small image costs a launch of its own; in a batch the images
// share it, and more sweeps than the batch's measured choice were found to
// cost more than they save.  The descriptors with their rounds go up from
// pinned staging (nothing waits for the copy: the staging is not written
// again before the call's final synchronisation).
int learned_plan_end_batch(Run& r, SchedulePlan& plan) {
    iwpp_workspace* ws = r.ws;
    const int n = (int)r.h.size();
    const LearnedLayout L = learned_layout(r, plan.tile);
    const int rw = learned_plan_wait(r);
    if (rw) return rw;
    const float* pred = (const float*)ws->host_plan;
    const int cap = std::max(1, plan.sweep_rounds);
    int rmax = 0;
    long long skipped = 0;
    for (int i = 0; i < n; ++i) {
        // at least the round already under way (its upward pass is still to come)
        r.h[i].rounds = std::max(1, std::min(cap, image_rounds(r.h[i], pred, L.base[i],
                                                               L.base[i + 1], L.tile)));
        rmax = std::max(rmax, r.h[i].rounds);
    }
    for (int i = 0; i < n; ++i) skipped += (long long)(rmax - r.h[i].rounds) * (L.base[i + 1] - L.base[i]);
    ImgDesc* h_desc = (ImgDesc*)round_up((uintptr_t)(pred + (size_t)L.total * 2) + (n + 1) * 4, 64);
    std::copy(r.h.begin(), r.h.end(), h_desc);
    HIP_TRY(hipMemcpyAsync(ws->descs, h_desc, sizeof(ImgDesc) * n, hipMemcpyHostToDevice, r.s));
    plan.sweep_rounds = rmax;
    r.st.schedule = IWPP_SCHED_LEARNED;
    r.st.sweeps_skipped_tiles = (int)std::min<long long>(skipped, 0x7FFFFFFF);
    return IWPP_OK;
}

// One image, tile t predicted to take k_t rounds (at least the first, which
// has already been launched).  Raster sweeps: the image runs R = the
// area-weighted mean of the k_t, rounded, in every tile.  A raster pass lasts
// as long as one wave takes down its column block, whatever the number of
// blocks that take part, so a round masked to part of the image costs about
// as much as a whole one (measured at 16384^2: a second round on a quarter of
// the tiles 0.45 ms, on all of them 0.45 ms): per-tile rounds buy nothing.
// Split sweeps (whose row passes skip the chunks of tiles that are off): R =
// the largest k_t that tiles of at least 1/64 of the image ask for (a tile
// asking for more gets R), and a tile with k_t < R skips the rounds from k_t
// on (tile masks).
int learned_plan_end_single(Run& r, SchedulePlan& plan) {
    iwpp_workspace* ws = r.ws;
    const ImgDesc& d = r.h[0];
    const int tile = plan.tile;
    const int ntx = (d.W + tile - 1) / tile, nty = (d.H + tile - 1) / tile, nt = ntx * nty;
    const int rw = learned_plan_wait(r);
    if (rw) return rw;
    const float* pred = (const float*)ws->host_plan;
    uint32_t* h_order = (uint32_t*)round_up((uintptr_t)(pred + (size_t)nt * 2) + 2 * 4 +
                                                sizeof(ImgDesc),
                                            64);  // tiles[nt], starts[nt+1]
    uint8_t* h_masks = (uint8_t*)(h_order + 2 * nt + 1);
    std::vector<int> k(nt);
    std::vector<double> area_at(6, 0.0);  // image area asking for exactly r rounds
    double total_area = 0, weighted = 0;
    for (int t = 0; t < nt; ++t) {
        const int tx = t % ntx, ty = t / ntx;
        const double area = (double)(std::min(d.W, (tx + 1) * tile) - tx * tile) *
                            (double)(std::min(d.H, (ty + 1) * tile) - ty * tile);
        k[t] = std::max(1, clamp_rounds(pred[(size_t)t * 2]));
        area_at[k[t]] += area;
        total_area += area;
        weighted += area * k[t];
    }
    int rounds = 4;
    if (!r.opt.split_sweeps) {
        rounds = std::max(1, (int)std::lround(weighted / std::max(1.0, total_area)));
    } else {
        while (rounds > 1 && area_at[rounds] < total_area / 64.0) {
            area_at[rounds - 1] += area_at[rounds];
            --rounds;
        }
    }
    for (int t = 0; t < nt; ++t) k[t] = r.opt.split_sweeps ? std::min(k[t], rounds) : rounds;
    plan.sweep_rounds = rounds;
    r.st.schedule = IWPP_SCHED_LEARNED;
    int skipped = 0;
    for (int t = 0; t < nt; ++t) skipped += std::max(0, rounds - k[t]);
    r.st.sweeps_skipped_tiles = skipped;
    if (nt == 1) return IWPP_OK;  // one tile: uniform rounds, no masks
    // The network's second output (tile priority) is not applied: seeding
    // the queue tile by tile in decreasing priority was measured at 32768^2
    // to cost more visits than index order (137K against 131K, +0.085 ms).
    if (skipped > 0) {  // per-round tile masks, only when some tile stops early
        const size_t nm = (size_t)rounds * nt;
        for (int rr = 0; rr < rounds; ++rr)
            for (int t = 0; t < nt; ++t) h_masks[(size_t)rr * nt + t] = k[t] > rr ? 1 : 0;
        HIP_TRY(grow(ws->tiles, ws->tiles_n, nm));
        HIP_TRY(hipMemcpyAsync(ws->tiles, h_masks, nm, hipMemcpyHostToDevice, r.s));
        r.round_masks = ws->tiles;
        r.sched_tile = tile;
        r.sched_ntx = ntx;
        r.sched_nt = nt;
    }
    return IWPP_OK;
}

// Seed class of the run's markers (autotuner key), from every 64th row.
int measure_seed_class(Run& r) {
    iwpp_workspace* ws = r.ws;
    int rc = ensure_ctl(ws);
    if (rc) return rc;
    HIP_TRY(grow(ws->cls, ws->cls_n, 2));
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(ws->cls), 0u, 4, r.s));
    int max_h = 0;
    for (const ImgDesc& d : r.h) max_h = std::max(max_h, d.H);
    HIP_TRY(launch_marker_class(r.ds, r.dtype, ws->cls, max_h, r.s));
    HIP_TRY(hipMemcpyAsync(ws->host_ctl, ws->cls, 2 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, r.s));
    HIP_TRY(hipStreamSynchronize(r.s));
    const unsigned long long* h = reinterpret_cast<const unsigned long long*>(ws->host_ctl);
    r.seed_class = seed_class(h[0], h[1]);
    return IWPP_OK;
}

// Plan, then the autotuner's measured choices for this workload class over
// the heuristic defaults (queue = AUTO, sweep_rounds < 0, col_block_rows = 0),
// then the images' descriptors (which depend on the column block height).
int plan_and_setup(Run& r, const iwpp_image* imgs, int n, int tile, int queue, int schedule,
                   SchedulePlan* plan) {
    int maxw = 0, maxh = 0;
    long long px = 0;
    for (int i = 0; i < n; ++i) {
        maxw = std::max(maxw, imgs[i].width);
        maxh = std::max(maxh, imgs[i].height);
        px += (long long)imgs[i].width * imgs[i].height;
    }
    if (plan_schedule(n == 1 ? imgs[0].width : maxw, n == 1 ? imgs[0].height : maxh, n, r.dtype,
                      r.conn, tile, queue, schedule, r.opt.sweep_rounds, plan))
        return fail(IWPP_ERR_ARG, "no schedule for this configuration");
    bool ready = false;
    if (!r.no_tune && tune_any() &&
        (queue == IWPP_QUEUE_AUTO || r.opt.sweep_rounds < 0 || r.opt.col_block_rows == 0)) {
        int rc = setup_images(r, imgs, n);
        if (rc) return rc;
        ready = true;
        if (r.seed_class < 0 && !r.capturing) {
            rc = measure_seed_class(r);
            if (rc) return rc;
        }
        TuneValue tv{};
        if (r.seed_class >= 0 &&
            tune_lookup(tune_key(r.dtype, r.conn, n, px, std::max(maxw, maxh), r.seed_class), &tv)) {
            if (queue == IWPP_QUEUE_AUTO) plan->queue = r.graph ? Q_MUTEX : tv.queue;
            if (r.opt.sweep_rounds < 0) plan->sweep_rounds = tv.sweep_rounds;
            if (r.opt.col_block_rows == 0 && tv.col_block_rows > 0) {
                r.opt.col_block_rows = tv.col_block_rows;
                // another block height: descriptors again
                if (r.h.empty() || r.h[0].R != tv.col_block_rows) ready = false;
            }
        }
    }
    r.queue = plan->queue;
    if (!ready) return setup_images(r, imgs, n);
    return IWPP_OK;
}

// Shared body of the single-image and batch entry points.  With graph = true
// only the launches are issued (nothing synchronises, no restart).
int reconstruct_run(Run& r, const iwpp_image* imgs, int n, int tile, int queue, int schedule,
                    const iwpp_options* opt) {
    iwpp_default_options(&r.opt);
    if (opt) r.opt = *opt;
    SchedulePlan plan;
    int rc = plan_and_setup(r, imgs, n, tile, queue, schedule, &plan);
    if (rc) return rc;
    r.st.queue = plan.queue;
    r.st.tile = plan.tile;
    r.st.schedule = IWPP_SCHED_PLAIN;
    r.st.col_block_rows = r.h.empty() ? 0 : r.h[0].R;
    r.st.seam_rows = r.seam;
    r.st.seed_class = r.seed_class;
    if (r.graph && r.queue != Q_MUTEX)
        return fail(IWPP_ERR_UNSUPPORTED, "graph capture needs the mutex queue variant");
    bool planning = false;
    // The learned plan costs about 30-50 us (tile features, network, and the
    // host planning between the first downward and upward column passes,
    // which the GPU waits for when those passes are short).  Measured on
    // 6144^2 and 8192^2 images, that is 5-10% of the whole reconstruction and
    // more than the best schedule saves over the plain rule, so a single image
    // below opt.learned_min_pixels keeps the plain rule; batches always plan.
    long long px_total = 0;
    for (int i = 0; i < n; ++i) px_total += (long long)imgs[i].width * imgs[i].height;
    if (schedule == IWPP_SCHED_LEARNED && !r.graph &&
        (n > 1 || px_total >= r.opt.learned_min_pixels)) {
        rc = learned_plan_begin(r, plan, &planning);
        if (rc) return rc;
    }
    if (r.graph) {
        rc = setup_queue(r);
        if (rc) return rc;
    }
    r.timed = r.timed && !r.graph;
    if (r.timed) {
        rc = ensure_events(r.ws);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(r.ws->ev[0], r.s));
    }
    if (planning) {
        r.plan_end = &plan;  // completed inside the first round (sweep_phase)
        rc = sweep_phase(r, 0, 1);
        if (rc) return rc;
        rc = sweep_phase(r, 1, plan.sweep_rounds);
    } else {
        r.seed_round = plan.sweep_rounds - 1;
        rc = sweep_phase(r, 0, plan.sweep_rounds);
    }
    if (rc) return rc;
    if (r.graph) return queue_launches(r);
    rc = queue_phase(r);
    if (rc) return rc;
    if (r.timed) {
        float a = 0;
        if (r.st.restarts == 0) {
            HIP_TRY(hipEventElapsedTime(&a, r.ws->ev[0], r.ws->ev[1]));
            const hipEvent_t* er = r.ws->ev_round.data();
            for (int k = 0; k < r.timed_rounds; ++k) {
                const hipEvent_t* e = er + kRoundEvents * k;
                const hipEvent_t end = k + 1 < r.timed_rounds ? e[kRoundEvents] : r.ws->ev[1];
                float rows = 0, cols = 0, fix0 = 0, fix1 = 0;
                HIP_TRY(hipEventElapsedTime(&rows, e[0], e[1]));
                if (r.opt.split_sweeps) {
                    HIP_TRY(hipEventElapsedTime(&cols, e[2], end));
                } else {  // raster: the fix-up passes between the events
                    HIP_TRY(hipEventElapsedTime(&fix0, e[1], e[2]));
                    HIP_TRY(hipEventElapsedTime(&cols, e[2], e[3]));
                    HIP_TRY(hipEventElapsedTime(&fix1, e[3], end));
                }
                r.st.ms_rows += rows;
                r.st.ms_cols += cols;
                r.st.ms_seams += fix0 + fix1;
            }
        }
        r.st.ms_sweep = a;
    }
    return IWPP_OK;
}

}  // namespace

extern "C" int iwpp_reconstruct_ex(void* marker, const void* mask, size_t pm, size_t pk, int w,
                                   int h, int dtype, int conn, int tile, int queue, int schedule,
                                   const iwpp_options* opt, iwpp_stats* stats,
                                   iwpp_workspace* ws, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    rc = check_common(dtype, conn, tile, queue, schedule);
    if (rc) return rc;
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");
    Run r{};
    r.dtype = dtype;
    r.conn = conn;
    r.ws = ws;
    HIP_TRY(call_stream(ws, stream, &r.s));
    r.timed = stats != nullptr;
    const iwpp_image img{marker, mask, pm, pk, w, h};
    rc = reconstruct_run(r, &img, 1, tile, queue, schedule, opt);
    if (rc) return rc;
    if (stats) *stats = r.st;
    return IWPP_OK;
}

extern "C" int iwpp_reconstruct(void* marker, const void* mask, size_t pm, size_t pk, int w,
                                int h, int dtype, int conn, int tile, int queue, int schedule,
                                iwpp_workspace* ws, void* stream) {
    return iwpp_reconstruct_ex(marker, mask, pm, pk, w, h, dtype, conn, tile, queue, schedule,
                               nullptr, nullptr, ws, stream);
}

extern "C" int iwpp_reconstruct_batch(const iwpp_image* imgs, int n, int dtype, int conn,
                                      int tile, int queue, int schedule,
                                      const iwpp_options* opt, iwpp_stats* stats,
                                      iwpp_workspace* ws, void* stream) {
    g_err.clear();
    if (!imgs || n <= 0) return fail(IWPP_ERR_ARG, "empty batch");
    for (int i = 0; i < n; ++i) {
        int rc = check_image(imgs[i].marker, imgs[i].mask, imgs[i].pitch_marker,
                             imgs[i].pitch_mask, imgs[i].width, imgs[i].height, dtype);
        if (rc) return rc;
    }
    int rc = check_common(dtype, conn, tile, queue, schedule);
    if (rc) return rc;
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");
    Run r{};
    r.dtype = dtype;
    r.conn = conn;
    r.ws = ws;
    HIP_TRY(call_stream(ws, stream, &r.s));
    r.timed = stats != nullptr;
    rc = reconstruct_run(r, imgs, n, tile, queue, schedule, opt);
    if (rc) return rc;
    if (stats) *stats = r.st;
    return IWPP_OK;
}

struct iwpp_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    iwpp_workspace* ws = nullptr;
    Run run;  // queue arguments and counters layout of the captured launch
};

extern "C" iwpp_graph* iwpp_batch_graph_create(const iwpp_image* imgs, int n, int dtype,
                                               int conn, int tile, int queue, int schedule,
                                               const iwpp_options* opt, iwpp_workspace* ws,
                                               void* stream) {
    g_err.clear();
    if (!imgs || n <= 0 || !ws) {
        fail(IWPP_ERR_ARG, "graph capture needs images and an explicit workspace");
        return nullptr;
    }
    for (int i = 0; i < n; ++i)
        if (check_image(imgs[i].marker, imgs[i].mask, imgs[i].pitch_marker, imgs[i].pitch_mask,
                        imgs[i].width, imgs[i].height, dtype))
            return nullptr;
    if (check_common(dtype, conn, tile, queue, schedule)) return nullptr;
    auto* g = new (std::nothrow) iwpp_graph();
    if (!g) return nullptr;
    hipStream_t s = (hipStream_t)stream;
    Run& r = g->run;
    r.dtype = dtype;
    r.conn = conn;
    r.ws = ws;
    r.s = s;
    r.graph = true;
    r.capturing = true;
    g->ws = ws;
    // Workspace growth, descriptor upload and every allocation happen during
    // a dry run of the set-up before capture; the capture then only records
    // kernel launches.
    {
        Run dry{};
        dry.dtype = dtype;
        dry.conn = conn;
        dry.ws = ws;
        dry.s = s;
        dry.graph = true;
        iwpp_default_options(&dry.opt);
        if (opt) dry.opt = *opt;
        SchedulePlan plan;
        if (plan_and_setup(dry, imgs, n, tile, queue, schedule, &plan)) {
            delete g;
            return nullptr;
        }
        if (plan.queue != Q_MUTEX) {
            fail(IWPP_ERR_UNSUPPORTED, "graph capture needs the mutex queue variant");
            delete g;
            return nullptr;
        }
        if (setup_queue(dry)) {
            delete g;
            return nullptr;
        }
        r.seed_class = dry.seed_class;  // the capture takes the same tuned choices
    }
    if (s == nullptr) {
        fail(IWPP_ERR_ARG, "graph capture needs a stream other than the NULL stream");
        delete g;
        return nullptr;
    }
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        fail(IWPP_ERR_HIP, "hipStreamBeginCapture failed");
        delete g;
        return nullptr;
    }
    const int rc = reconstruct_run(r, imgs, n, tile, queue, schedule, opt);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc || e != hipSuccess || !graph) {
        if (!rc) fail(IWPP_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        if (graph) (void)hipGraphDestroy(graph);
        delete g;
        return nullptr;
    }
    g->graph = graph;
    if (hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        fail(IWPP_ERR_HIP, "hipGraphInstantiate failed");
        (void)hipGraphDestroy(graph);
        delete g;
        return nullptr;
    }
    return g;
}

extern "C" int iwpp_batch_graph_launch(iwpp_graph* g, void* stream) {
    g_err.clear();
    if (!g || !g->exec) return fail(IWPP_ERR_ARG, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_batch_graph_stats(iwpp_graph* g, iwpp_stats* stats, void* stream) {
    g_err.clear();
    if (!g) return fail(IWPP_ERR_ARG, "null graph");
    hipStream_t s = (hipStream_t)stream;
    iwpp_workspace* ws = g->ws;
    HIP_TRY(hipMemcpyAsync(ws->host_ctl, ws->ctl, (size_t)C_NSLOTS * kCtlStride * 4,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    Run& r = g->run;
    r.st = iwpp_stats{};
    r.st.queue = r.queue;
    const int rc = read_counters(r);
    if (rc < 0) return rc;
    if (rc > 0) return fail(IWPP_ERR_UNSUPPORTED, "queue overflow inside a graph replay");
    if (stats) *stats = r.st;
    return IWPP_OK;
}

extern "C" void iwpp_batch_graph_destroy(iwpp_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

extern "C" int iwpp_autotune(const iwpp_image* imgs, int n, int dtype, int conn, int trials,
                             const iwpp_options* opt, iwpp_tuning* best, iwpp_workspace* ws,
                             void* stream) {
    g_err.clear();
    if (!imgs || n <= 0) return fail(IWPP_ERR_ARG, "empty batch");
    int rc = check_common(dtype, conn, 0, IWPP_QUEUE_AUTO, IWPP_SCHED_PLAIN);
    if (rc) return rc;
    long long px = 0;
    int maxside = 0;
    size_t snap_bytes = 0;
    for (int i = 0; i < n; ++i) {
        rc = check_image(imgs[i].marker, imgs[i].mask, imgs[i].pitch_marker, imgs[i].pitch_mask,
                         imgs[i].width, imgs[i].height, dtype);
        if (rc) return rc;
        px += (long long)imgs[i].width * imgs[i].height;
        maxside = std::max(maxside, std::max(imgs[i].width, imgs[i].height));
        snap_bytes += (size_t)imgs[i].width * dtype * imgs[i].height;
    }
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");
    hipStream_t s = nullptr;
    HIP_TRY(call_stream(ws, stream, &s));
    // snapshot of the input markers (packed rows), restored before every trial
    uint8_t* snap = nullptr;
    HIP_TRY(hipMalloc((void**)&snap, std::max<size_t>(snap_bytes, 1)));
    auto copy_all = [&](bool to_snap) -> hipError_t {
        size_t off = 0;
        for (int i = 0; i < n; ++i) {
            const size_t row = (size_t)imgs[i].width * dtype;
            hipError_t e = to_snap
                ? hipMemcpy2DAsync(snap + off, row, imgs[i].marker, imgs[i].pitch_marker, row,
                                   imgs[i].height, hipMemcpyDeviceToDevice, s)
                : hipMemcpy2DAsync(imgs[i].marker, imgs[i].pitch_marker, snap + off, row, row,
                                   imgs[i].height, hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return e;
            off += row * imgs[i].height;
        }
        return hipSuccess;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = copy_all(true);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    // the workload's seed class, for the cache key
    int sclass = 0;
    if (he == hipSuccess) {
        Run r0{};
        r0.dtype = dtype;
        r0.conn = conn;
        r0.ws = ws;
        r0.s = s;
        iwpp_default_options(&r0.opt);
        rc = setup_images(r0, imgs, n);
        if (rc == IWPP_OK) rc = measure_seed_class(r0);
        sclass = r0.seed_class;
    }
    // One trial configuration: best of `trials` timed reconstructions.
    auto trial = [&](int queue, int sw, int rows) -> float {
        float best_ms = 1e30f;
        for (int t = 0; t < std::max(1, trials) && rc == IWPP_OK && he == hipSuccess; ++t) {
            if ((he = copy_all(false)) != hipSuccess) break;
            if ((he = hipEventRecord(e0, s)) != hipSuccess) break;
            Run r{};
            r.dtype = dtype;
            r.conn = conn;
            r.ws = ws;
            r.s = s;
            r.no_tune = true;
            iwpp_options o;
            iwpp_default_options(&o);
            if (opt) o = *opt;
            o.sweep_rounds = sw;
            o.col_block_rows = rows;
            rc = reconstruct_run(r, imgs, n, 0, queue, IWPP_SCHED_PLAIN, &o);
            if (rc) break;
            if ((he = hipEventRecord(e1, s)) != hipSuccess) break;
            if ((he = hipEventSynchronize(e1)) != hipSuccess) break;
            float ms = 0;
            if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            best_ms = std::min(best_ms, ms);
        }
        return best_ms;
    };
    // Stage 1: queue variant x sweep rounds, column blocks from the size
    // rule; stage 2: the column block height (the sweep passes' tile: how far
    // a pass carries a rise before the queue takes over) for the winner.
    iwpp_tuning top{IWPP_QUEUE_MUTEX, 0, 0, 1e30f, 0, sclass};
    const int queues[3] = {IWPP_QUEUE_MUTEX, IWPP_QUEUE_DQ, IWPP_QUEUE_UQ};
    for (int qi = 0; qi < 3 && he == hipSuccess && rc == IWPP_OK; ++qi) {
        for (int sw = 0; sw <= 4 && he == hipSuccess && rc == IWPP_OK; ++sw) {
            const float ms = trial(queues[qi], sw, 0);
            if (ms < top.ms) top = iwpp_tuning{queues[qi], sw, 0, ms, 0, sclass};
        }
    }
    if (top.sweep_rounds > 0) {
        const int rows_c[4] = {128, 256, 512, 1024};
        for (int k = 0; k < 4 && he == hipSuccess && rc == IWPP_OK; ++k) {
            const float ms = trial(top.queue, top.sweep_rounds, rows_c[k]);
            if (ms < 0.98f * top.ms) {  // a clear win over the size rule's height
                top.ms = ms;
                top.col_block_rows = rows_c[k];
            }
        }
    }
    if (he == hipSuccess) he = copy_all(false);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(snap);
    if (rc) return rc;
    if (he != hipSuccess) return fail(IWPP_ERR_HIP, "autotune: %s", hipGetErrorString(he));
    SchedulePlan plan;
    plan_schedule(maxside, maxside, n, dtype, conn, 0, top.queue, IWPP_SCHED_PLAIN,
                  top.sweep_rounds, &plan);
    top.tile = plan.tile;
    tune_store(tune_key(dtype, conn, n, px, maxside, sclass),
               TuneValue{top.queue, top.sweep_rounds, top.col_block_rows});
    if (best) *best = top;
    return IWPP_OK;
}

extern "C" void iwpp_autotune_clear(void) { tune_clear(); }

extern "C" int iwpp_sched_set_model(const float* weights, int n) {
    g_err.clear();
    static_assert(kMlpParams == IWPP_MLP_PARAMS && kMlpIn == IWPP_MLP_FEATURES, "MLP shape");
    if (weights && n != kMlpParams) return fail(IWPP_ERR_ARG, "model needs %d weights", kMlpParams);
    std::vector<uint16_t> bf;
    if (weights) {
        bf.resize(kMlpParams);
        for (int i = 0; i < kMlpParams; ++i) {
            if (!std::isfinite(weights[i])) return fail(IWPP_ERR_ARG, "weight %d not finite", i);
            bf[i] = f32_to_bf16(weights[i]);
        }
    }
    std::lock_guard<std::mutex> lk(g_model_mu);
    g_model.swap(bf);
    ++g_model_version;
    return IWPP_OK;
}

extern "C" int iwpp_tile_features(const void* marker, const void* mask, size_t pm, size_t pk,
                                  int w, int h, int dtype, int conn, int tile, int sub,
                                  float* raw, float* feat, iwpp_workspace* ws, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (tile < kMP || tile % kMP) return fail(IWPP_ERR_ARG, "tile must be a multiple of 64");
    if (sub < 1) return fail(IWPP_ERR_ARG, "sub must be >= 1");
    if (!raw || !feat) return fail(IWPP_ERR_ARG, "null output");
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");
    const size_t nt = (size_t)((w + tile - 1) / tile) * ((h + tile - 1) / tile);
    HIP_TRY(grow(ws->feat_part, ws->feat_part_n, nt * features_slabs(tile, sub, dtype) * 16));
    HIP_TRY(launch_tile_features(Image{(uint8_t*)marker, (const uint8_t*)mask, pm, pk, w, h},
                                 dtype, conn, tile, sub, ws->feat_part, raw, feat,
                                 (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_mlp_forward(const void* params, int bf16, const float* x, int n, float* y,
                                float* acts, void* stream) {
    g_err.clear();
    if (!params || n < 0 || (n > 0 && (!x || !y))) return fail(IWPP_ERR_ARG, "bad MLP arguments");
    HIP_TRY(launch_mlp_forward(params, bf16 != 0, x, n, y, acts, (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_mlp_backward(const void* params, int bf16, const float* x, const float* acts,
                                 const float* y, const float* t, int n, float lw0, float lw1,
                                 float* work, float* grad, void* stream) {
    g_err.clear();
    if (!params || !grad || n <= 0 || !x || !acts || !y || !t || !work)
        return fail(IWPP_ERR_ARG, "bad MLP arguments");
    float* d3 = work;
    float* d2 = d3 + (size_t)n * 2;
    float* d1 = d2 + (size_t)n * 64;
    HIP_TRY(launch_mlp_backward(params, bf16 != 0, x, acts, y, t, n, lw0, lw1, d3, d2, d1, grad,
                                (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_mlp_adam(float* w, float* m, float* v, const float* grad, uint16_t* wbf,
                             float lr, int step, void* stream) {
    g_err.clear();
    if (!w || !m || !v || !grad || !wbf || step < 1) return fail(IWPP_ERR_ARG, "bad Adam arguments");
    HIP_TRY(launch_adam(w, m, v, grad, wbf, lr, step, (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_clip(void* marker, const void* mask, size_t pm, size_t pk, int w, int h,
                         int dtype, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    HIP_TRY(launch_clip(Image{(uint8_t*)marker, (const uint8_t*)mask, pm, pk, w, h}, dtype,
                        (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_sweep_rows(void* marker, const void* mask, size_t pm, size_t pk, int w,
                               int h, int dtype, int* changed, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    const TileMask none{nullptr, kMP, kMP, 1};
    HIP_TRY(launch_row_pass(Image{(uint8_t*)marker, (const uint8_t*)mask, pm, pk, w, h}, dtype,
                            false, none, (uint32_t*)changed, (hipStream_t)stream));
    return IWPP_OK;
}

namespace {
// One out-of-place blocked pass (column or raster) of a single image.
int blocked_pass(const void* src, void* dst, const void* mask, size_t ps, size_t pd, size_t pk,
                 int w, int h, int dtype, int conn, int dir, int block_rows, int mode,
                 int* changed, void* stream) {
    g_err.clear();
    int rc = check_image(src, mask, ps, pk, w, h, dtype);
    if (rc) return rc;
    rc = check_image(dst, mask, pd, pk, w, h, dtype);
    if (rc) return rc;
    if (src == dst) return fail(IWPP_ERR_ARG, "the pass is out of place: src == dst");
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (dir != 1 && dir != -1) return fail(IWPP_ERR_ARG, "dir must be +1 or -1");
    const bool wide = col_wide(src, dst, mask, ps, pd, pk, w, dtype);
    Descs ds{};
    ds.n = 1;
    ImgDesc& d = ds.one;
    d.J = (uint8_t*)src;  // read J, write P
    d.P = (uint8_t*)dst;
    d.I = (const uint8_t*)mask;
    d.pj = ps;
    d.pp = pd;
    d.pi = pk;
    d.W = w;
    d.H = h;
    d.strips = col_strips(w, dtype, wide);
    d.R = block_rows > 0 ? block_rows : col_auto_rows(w, h, dtype, wide);
    d.tm = TileMask{nullptr, kMP, kMP, 1};
    ds.total = d.strips * ((h + d.R - 1) / d.R);
    HIP_TRY(launch_col_pass_batch(ds, dtype, conn, dir, mode, wide, (uint32_t*)changed,
                                  (hipStream_t)stream));
    return IWPP_OK;
}
}  // namespace

extern "C" int iwpp_sweep_cols(const void* src, void* dst, const void* mask, size_t ps,
                               size_t pd, size_t pk, int w, int h, int dtype, int conn, int dir,
                               int block_rows, int* changed, void* stream) {
    return blocked_pass(src, dst, mask, ps, pd, pk, w, h, dtype, conn, dir, block_rows,
                        COL_J_TO_P, changed, stream);
}

extern "C" int iwpp_sweep_raster(const void* src, void* dst, const void* mask, size_t ps,
                                 size_t pd, size_t pk, int w, int h, int dtype, int conn,
                                 int dir, int block_rows, int* changed, void* stream) {
    return blocked_pass(src, dst, mask, ps, pd, pk, w, h, dtype, conn, dir, block_rows,
                        COL_RASTER_J_TO_P, changed, stream);
}

extern "C" int iwpp_sweep_cols_block_width(const void* src, const void* dst, const void* mask,
                                           size_t ps, size_t pd, size_t pk, int w, int dtype) {
    if (dtype != IWPP_U8 && dtype != IWPP_U16) return fail(IWPP_ERR_ARG, "bad dtype");
    return col_block_width(src, dst, mask, ps, pd, pk, w, dtype);
}

extern "C" int iwpp_unstable_megapixels(const void* marker, const void* mask, size_t pm,
                                        size_t pk, int w, int h, int dtype, int conn,
                                        uint8_t* flags, int* count, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (!flags) return fail(IWPP_ERR_ARG, "null flags");
    const bool wide = col_wide(marker, marker, mask, pm, pm, pk, w, dtype);
    Descs ds{};
    ds.n = 1;
    ImgDesc& d = ds.one;
    d.J = (uint8_t*)marker;
    d.I = (const uint8_t*)mask;
    d.pj = pm;
    d.pi = pk;
    d.W = w;
    d.H = h;
    d.nmx = (w + kMP - 1) / kMP;
    d.nmy = (h + kMP - 1) / kMP;
    d.strips = col_strips(w, dtype, wide);
    ds.total = d.strips * d.nmy;
    HIP_TRY(launch_unstable_batch(ds, dtype, conn, wide, flags, (uint32_t*)count,
                                  (hipStream_t)stream));
    return IWPP_OK;
}

extern "C" int iwpp_sweep_round_seed(const iwpp_image* imgs, int n, int dtype, int conn,
                                     int block_rows, const iwpp_options* opt, uint8_t* flags,
                                     iwpp_workspace* ws, void* stream) {
    g_err.clear();
    if (!imgs || n <= 0) return fail(IWPP_ERR_ARG, "empty batch");
    if (!flags) return fail(IWPP_ERR_ARG, "null flags");
    for (int i = 0; i < n; ++i) {
        const int rc = check_image(imgs[i].marker, imgs[i].mask, imgs[i].pitch_marker,
                                   imgs[i].pitch_mask, imgs[i].width, imgs[i].height, dtype);
        if (rc) return rc;
    }
    int rc = check_common(dtype, conn, 0, IWPP_QUEUE_MUTEX, IWPP_SCHED_PLAIN);
    if (rc) return rc;
    if (block_rows < 0 || block_rows % 32) return fail(IWPP_ERR_ARG, "block_rows must be 0 or a multiple of 32");
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");
    Run r{};
    r.dtype = dtype;
    r.conn = conn;
    r.ws = ws;
    r.queue = Q_MUTEX;
    r.no_tune = true;
    iwpp_default_options(&r.opt);
    if (opt) {
        r.opt.split_sweeps = opt->split_sweeps ? 1 : 0;
        r.opt.seam_rows = opt->seam_rows;
    }
    r.opt.col_block_rows = block_rows;
    HIP_TRY(call_stream(ws, stream, &r.s));
    rc = setup_images(r, imgs, n);
    if (rc) return rc;
    r.seed_round = 0;  // the upward pass of this one round sets the flags
    rc = sweep_phase(r, 0, 1);
    if (rc) return rc;
    if (!r.flags_ready) return fail(IWPP_ERR_UNSUPPORTED, "the seeding pass did not run");
    HIP_TRY(launch_seed_edges(with_total(r.ds, r.total_se), r.dtype, r.conn, r.wide, ws->edges,
                              ws->flags, r.s));
    HIP_TRY(hipMemcpyAsync(flags, ws->flags, (size_t)r.total_mp, hipMemcpyDeviceToDevice, r.s));
    if (!stream) HIP_TRY(hipStreamSynchronize(r.s));
    return IWPP_OK;
}

// ---------------------------------------------- allocation-free entry point --
// iwpp_reconstruct_noalloc runs the single-image, plain-schedule, mutex-queue
// reconstruction of iwpp_reconstruct_ex with every device buffer it needs
// beyond the caller's images carved out of one static device array, so that
// it never allocates.  Layout (bytes from the array's start):
//   control block | megapixel states (u32) | seeding flags (u8) | queue ring (u64)
// Seeding takes k_unstable_scan's flags (the fused form of the last column
// pass needs a strip-edge buffer that grows with the image's height).
namespace {
constexpr size_t kNaMaxMP = IWPP_NOALLOC_MAX_MP;
constexpr size_t kNaRing = 524288;  // ring entries: next_pow2(kNaMaxMP + 16384)
constexpr size_t kNaCtlBytes = 4096;
constexpr size_t kNaStOff = kNaCtlBytes;
constexpr size_t kNaFlagsOff = kNaStOff + kNaMaxMP * 4;
constexpr size_t kNaRingOff = kNaFlagsOff + kNaMaxMP;
constexpr size_t kNaBytes = kNaRingOff + kNaRing * 8;
static_assert((size_t)C_NSLOTS * kCtlStride * 4 <= kNaCtlBytes, "control block fits");
static_assert(kNaRing >= kNaMaxMP + 16384 && (kNaRing & (kNaRing - 1)) == 0, "ring size");
static_assert(kNaRingOff % 256 == 0, "ring alignment");
}  // namespace

// One copy per GPU, held until the process ends.  In host code this name
// only stands for the symbol: its device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};

// The current device's address of `scratch` (looked up on the first call
// with that device current).
int noalloc_scratch(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return fail(IWPP_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(IWPP_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_na_ptr[dev] = p;
    }
    *out = g_na_ptr[dev];
    return IWPP_OK;
}

struct NoAllocScope {
    NoAllocScope() { t_no_alloc = true; }
    ~NoAllocScope() { t_no_alloc = false; }
};
}  // namespace

extern "C" int iwpp_reconstruct_noalloc(void* marker, const void* mask, size_t pm, size_t pk,
                                        int w, int h, int dtype, int conn,
                                        const iwpp_options* opt, void* stream) {
    g_err.clear();
    int rc = check_image(marker, mask, pm, pk, w, h, dtype);
    if (rc) return rc;
    rc = check_common(dtype, conn, 0, IWPP_QUEUE_MUTEX, IWPP_SCHED_PLAIN);
    if (rc) return rc;
    iwpp_options o;
    iwpp_default_options(&o);
    if (opt) o = *opt;
    if (o.queue_capacity != 0)
        return fail(IWPP_ERR_ARG, "queue_capacity must be 0: the allocation-free entry point "
                                  "sizes its queue itself");
    const size_t nmp = (size_t)((w + kMP - 1) / kMP) * (size_t)((h + kMP - 1) / kMP);
    if (nmp > kNaMaxMP)
        return fail(IWPP_ERR_ARG, "image of %zu megapixels (64 x 64 pixels each): the "
                                  "allocation-free entry point takes at most %zu",
                    nmp, kNaMaxMP);
    void* base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    // A workspace whose buffers are views of the scratch, each at its full
    // size, so that nothing grows; never passed to iwpp_workspace_destroy.
    iwpp_workspace fx;
    fx.device = -1;
    fx.ctl = static_cast<uint32_t*>(base);
    fx.ctl_n = kNaCtlBytes / 4;
    fx.st = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + kNaStOff);
    fx.st_n = kNaMaxMP;
    fx.flags = reinterpret_cast<uint8_t*>(static_cast<char*>(base) + kNaFlagsOff);
    fx.flags_n = kNaMaxMP;
    fx.ring = reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + kNaRingOff);
    fx.ring_n = kNaRing;
    alignas(16) uint32_t host_ctl[(size_t)C_NSLOTS * kCtlStride];  // counters read back here
    fx.host_ctl = host_ctl;
    Run r{};
    r.dtype = dtype;
    r.conn = conn;
    r.ws = &fx;
    r.s = (hipStream_t)stream;  // NULL: the NULL stream itself (the mutex kernel is an
                                // ordinary launch)
    r.timed = false;
    r.scan_seed = true;
    r.no_tune = true;  // (measuring the seed class would need scratch of its own)
    const iwpp_image img{marker, mask, pm, pk, w, h};
    NoAllocScope guard;
    rc = reconstruct_run(r, &img, 1, 0, IWPP_QUEUE_MUTEX, IWPP_SCHED_PLAIN, &o);
    if (rc == IWPP_ERR_HIP && std::strstr(g_err.c_str(), "grow("))
        return fail(IWPP_ERR_UNSUPPORTED, "the allocation-free call needed device memory "
                                          "beyond its static scratch");
    return rc;
}

extern "C" const char* iwpp_last_error(void) { return g_err.c_str(); }

extern "C" void MorphologicalReconstruction(unsigned char* marker, unsigned char* mask,
                                            size_t pitchMarker, size_t pitchMask, int size) {
    int queue = IWPP_QUEUE_UQ;
    if (const char* q = std::getenv("IWPP_QUEUE")) {
        if (!strcmp(q, "mutex")) queue = IWPP_QUEUE_MUTEX;
        else if (!strcmp(q, "dq")) queue = IWPP_QUEUE_DQ;
        else if (!strcmp(q, "uq")) queue = IWPP_QUEUE_UQ;
    }
    int rc = iwpp_reconstruct(marker, mask, pitchMarker, pitchMask, size, size, IWPP_U8, 4, 0,
                              queue, IWPP_SCHED_PLAIN, nullptr, nullptr);
    if (rc == IWPP_OK && hipDeviceSynchronize() != hipSuccess) rc = IWPP_ERR_HIP;
    if (rc != IWPP_OK) fprintf(stderr, "MorphologicalReconstruction: %s\n", iwpp_last_error());
}
