// Host driver and C ABI (include/iwpp.h).  Orchestrates, on one HIP stream:
//   clip (folded into the first row pass) -> sweep rounds (row pass in place,
//   column pass down into a scratch plane, column pass up back) -> seeding
//   (unstable megapixels compacted into the queue) -> queue phase (mutex /
//   DQ / UQ) -> overflow restart from the partially reconstructed marker.
// The restart mirrors the reference's do { ... } while (memLeak) loop
// (IWPP MP/src/MorphologicalReconstruction.cu:785-806,
//  IWPP MP DQ/src/MorphologicalReconstruction.cu:1242-1323): the marker only
// rises and never passes the reconstruction, so any partial marker is a valid
// restart point.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "iwpp.h"
#include "iwpp_internal.h"
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

template <class P>
hipError_t grow(P*& p, size_t& have, size_t need_elems) {
    if (p && have >= need_elems) return hipSuccess;
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
    uint8_t* plane = nullptr;   // scratch plane for the out-of-place column passes
    size_t plane_n = 0;
    uint32_t* st = nullptr;     // per-megapixel state
    size_t st_n = 0;
    uint8_t* flags = nullptr;   // per-megapixel unstable flags
    size_t flags_n = 0;
    uint32_t* qa = nullptr;
    size_t qa_n = 0;
    uint32_t* qb = nullptr;
    size_t qb_n = 0;
    unsigned long long* ring = nullptr;  // mutex ring
    size_t ring_n = 0;
    uint32_t* ctl = nullptr;    // control block
    size_t ctl_n = 0;
    uint32_t* host_ctl = nullptr;  // pinned copy of the control block
    uint8_t* tiles = nullptr;   // per-tile active flags (schedule)
    size_t tiles_n = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // phase timing
};

extern "C" void iwpp_default_options(iwpp_options* o) {
    if (!o) return;
    o->sweep_rounds = -1;
    o->col_block_rows = 0;
    o->max_local_iters = 8;
    o->queue_capacity = 0;
    o->workers_per_cu = 0;
    o->timeout_s = 60.0;
    o->trace_phases = 0;
}

extern "C" iwpp_workspace* iwpp_workspace_create(size_t max_pixels, int max_dtype) {
    auto* ws = new (std::nothrow) iwpp_workspace();
    if (!ws) return nullptr;
    if (hipGetDevice(&ws->device) != hipSuccess) {
        delete ws;
        return nullptr;
    }
    if (max_pixels && max_dtype >= 1 && max_dtype <= 2) {
        if (grow(ws->plane, ws->plane_n, max_pixels * (size_t)max_dtype + 4096) != hipSuccess) {
            delete ws;
            return nullptr;
        }
    }
    return ws;
}

extern "C" void iwpp_workspace_destroy(iwpp_workspace* ws) {
    if (!ws) return;
    (void)hipFree(ws->plane);
    (void)hipFree(ws->st);
    (void)hipFree(ws->flags);
    (void)hipFree(ws->qa);
    (void)hipFree(ws->qb);
    (void)hipFree(ws->ring);
    (void)hipFree(ws->ctl);
    (void)hipFree(ws->tiles);
    if (ws->host_ctl) (void)hipHostFree(ws->host_ctl);
    for (hipEvent_t e : ws->ev)
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

struct Run {
    bool timed;  // record phase times with HIP events
    Image im;
    int dtype, conn, tile, queue, schedule;
    iwpp_options opt;
    iwpp_stats st;
    iwpp_workspace* ws;
    hipStream_t s;
    int nmx, nmy, nmp;
    size_t plane_pitch;
};

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

// Sweep phase: `rounds` rounds of row pass + column pass down + column pass up.
int sweep_phase(Run& r, int rounds, const TileMask& tm) {
    const Image& im = r.im;
    if (rounds <= 0) {
        HIP_TRY(launch_clip(im, r.dtype, r.s));
        return IWPP_OK;
    }
    const size_t need = r.plane_pitch * (size_t)im.H + 256;
    HIP_TRY(grow(r.ws->plane, r.ws->plane_n, need));
    uint8_t* plane = (uint8_t*)round_up((uintptr_t)r.ws->plane, 256);
    for (int k = 0; k < rounds; ++k) {
        HIP_TRY(launch_row_pass(im, r.dtype, k == 0, tm, nullptr, r.s));
        HIP_TRY(launch_col_pass(im.J, plane, im.I, im.pj, r.plane_pitch, im.pi, im.W, im.H,
                                r.dtype, r.conn, +1, r.opt.col_block_rows, tm, nullptr, r.s));
        HIP_TRY(launch_col_pass(plane, im.J, im.I, r.plane_pitch, im.pj, im.pi, im.W, im.H,
                                r.dtype, r.conn, -1, r.opt.col_block_rows, tm, nullptr, r.s));
    }
    r.st.sweep_rounds += rounds;
    return IWPP_OK;
}

// Seeding + queue phase, restarted while a queue overflowed.
int queue_phase(Run& r) {
    iwpp_workspace* ws = r.ws;
    const int kind = r.queue;
    const size_t nmp = (size_t)r.nmp;
    HIP_TRY(grow(ws->st, ws->st_n, nmp));
    HIP_TRY(grow(ws->flags, ws->flags_n, nmp));
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

    QueueArgs q{};
    q.ctl = ws->ctl;
    q.st = ws->st;
    q.qa = ws->qa;
    q.qb = ws->qb;
    q.ring = ws->ring;
    q.cap = (uint32_t)cap;
    q.nmx = r.nmx;
    q.nmy = r.nmy;
    q.max_iters = std::max(1, r.opt.max_local_iters);
    q.max_rounds = 1 << 30;
    q.timeout_ticks = timeout_ticks(r.opt.timeout_s > 0 ? r.opt.timeout_s : 60.0);
    q.trace = r.opt.trace_phases ? 1 : 0;
    const double us_per_tick = 1e6 / ((double)timeout_ticks(1.0));

    const int max_restarts = 100000;
    for (int attempt = 0;; ++attempt) {
        if (r.timed) HIP_TRY(hipEventRecord(ws->ev[1], r.s));
        HIP_TRY(launch_fill_u32(ws->ctl, 0u, (size_t)C_NSLOTS * kCtlStride, r.s));
        HIP_TRY(launch_fill_u32(ws->st, 0u, nmp, r.s));
        if (kind == Q_MUTEX) HIP_TRY(launch_ring_init(ws->ring, cap, r.s));
        HIP_TRY(launch_unstable(r.im, r.dtype, r.conn, ws->flags, nullptr, r.s));
        HIP_TRY(launch_seed_queue(ws->flags, r.nmp, kind, q, r.s));
        HIP_TRY(launch_prepare_queue(q, r.s));
        if (r.timed) HIP_TRY(hipEventRecord(ws->ev[2], r.s));
        HIP_TRY(launch_queue_phase(r.im, r.dtype, r.conn, kind, q, r.opt.workers_per_cu, r.s));
        if (r.timed) HIP_TRY(hipEventRecord(ws->ev[3], r.s));
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
        const uint32_t* h = ws->host_ctl;
        r.st.seeded += h[C_TAIL * kCtlStride] + 0ll;
        r.st.tile_visits += (long long)(*(const unsigned long long*)(h + C_VISITS_LO * kCtlStride));
        r.st.queue_rounds += (int)h[C_ROUND * kCtlStride];
        if (q.trace) {
            auto ph = [&](int slot) {
                return (double)(*(const unsigned long long*)(h + slot * kCtlStride)) * us_per_tick;
            };
            r.st.us_wait += ph(C_PH_POP);
            r.st.us_fetch += ph(C_PH_FETCH);
            r.st.us_converge += ph(C_PH_CONVERGE);
            r.st.us_store += ph(C_PH_STORE);
            r.st.us_notify += ph(C_PH_NOTIFY);
        }
        if (h[C_TIMEOUT * kCtlStride])
            return fail(IWPP_ERR_TIMEOUT, "device-side queue wait timed out");
        if (!h[C_OVERFLOW * kCtlStride]) break;
        r.st.restarts++;
        if (attempt >= max_restarts) return fail(IWPP_ERR_TIMEOUT, "too many queue restarts");
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
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (tile < 0 || tile % kMP) return fail(IWPP_ERR_ARG, "tile must be 0 or a multiple of 64");
    if (queue < IWPP_QUEUE_MUTEX || queue > IWPP_QUEUE_AUTO)
        return fail(IWPP_ERR_ARG, "bad queue variant %d", queue);
    if (schedule != IWPP_SCHED_PLAIN && schedule != IWPP_SCHED_LEARNED)
        return fail(IWPP_ERR_ARG, "bad schedule %d", schedule);
    if (!ws) ws = default_ws();
    if (!ws) return fail(IWPP_ERR_HIP, "no HIP device");

    Run r{};
    r.im = Image{(uint8_t*)marker, (const uint8_t*)mask, pm, pk, w, h};
    r.dtype = dtype;
    r.conn = conn;
    r.ws = ws;
    r.s = (hipStream_t)stream;
    iwpp_default_options(&r.opt);
    if (opt) r.opt = *opt;
    r.nmx = (w + kMP - 1) / kMP;
    r.nmy = (h + kMP - 1) / kMP;
    r.nmp = r.nmx * r.nmy;
    r.plane_pitch = round_up((size_t)w * dtype, 256);

    SchedulePlan plan;
    rc = plan_schedule(w, h, dtype, conn, tile, queue, schedule, r.opt.sweep_rounds, &plan);
    if (rc) return fail(IWPP_ERR_ARG, "no schedule for this configuration");
    r.tile = plan.tile;
    r.queue = plan.queue;
    r.schedule = schedule;
    r.st.queue = plan.queue;
    r.st.tile = plan.tile;

    r.timed = stats != nullptr;
    if (r.timed) {
        rc = ensure_events(ws);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ws->ev[0], r.s));
    }
    const TileMask none{nullptr, kMP, kMP, 1};
    rc = sweep_phase(r, plan.sweep_rounds, none);
    if (rc) return rc;
    rc = queue_phase(r);
    if (rc) return rc;
    if (r.timed) {
        // ev[0] -> first ev[1] would be exact; the seeding record of the last
        // attempt is the upper bound of the sweep phase when no restart happened.
        float a = 0;
        if (r.st.restarts == 0) HIP_TRY(hipEventElapsedTime(&a, ws->ev[0], ws->ev[1]));
        r.st.ms_sweep = a;
    }
    if (stats) *stats = r.st;
    return IWPP_OK;
}

extern "C" int iwpp_reconstruct(void* marker, const void* mask, size_t pm, size_t pk, int w,
                                int h, int dtype, int conn, int tile, int queue, int schedule,
                                iwpp_workspace* ws, void* stream) {
    return iwpp_reconstruct_ex(marker, mask, pm, pk, w, h, dtype, conn, tile, queue, schedule,
                               nullptr, nullptr, ws, stream);
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

extern "C" int iwpp_sweep_cols(const void* src, void* dst, const void* mask, size_t ps,
                               size_t pd, size_t pk, int w, int h, int dtype, int conn, int dir,
                               int block_rows, int* changed, void* stream) {
    g_err.clear();
    int rc = check_image(src, mask, ps, pk, w, h, dtype);
    if (rc) return rc;
    rc = check_image(dst, mask, pd, pk, w, h, dtype);
    if (rc) return rc;
    if (src == dst) return fail(IWPP_ERR_ARG, "column pass is out of place: src == dst");
    if (conn != 4 && conn != 8) return fail(IWPP_ERR_ARG, "connectivity must be 4 or 8");
    if (dir != 1 && dir != -1) return fail(IWPP_ERR_ARG, "dir must be +1 or -1");
    const TileMask none{nullptr, kMP, kMP, 1};
    HIP_TRY(launch_col_pass((const uint8_t*)src, (uint8_t*)dst, (const uint8_t*)mask, ps, pd, pk,
                            w, h, dtype, conn, dir, block_rows, none, (uint32_t*)changed,
                            (hipStream_t)stream));
    return IWPP_OK;
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
    HIP_TRY(launch_unstable(Image{(uint8_t*)marker, (const uint8_t*)mask, pm, pk, w, h}, dtype,
                            conn, flags, (uint32_t*)count, (hipStream_t)stream));
    return IWPP_OK;
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
