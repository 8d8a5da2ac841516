// iwpp_recon: command-line reconstruction, the counterpart of the reference's
// test program (IWPP MP/src/main.cu:11-69, IWPP MP DQ/src/main.cu:128-240):
// read marker and mask, upload them to pitched device memory, reconstruct,
// print the time, write the result.  Files are PGM (P5/P2, 8- or 16-bit,
// include/iwpp_io.h) instead of the reference's OpenCV formats.
//
//   iwpp_recon MARKER.pgm MASK.pgm OUT.pgm [options]
//     --conn 4|8            connectivity (default 4, the reference's)
//     --queue mutex|dq|uq|auto   queue variant (default auto)
//     --tile N              scheduling tile, a multiple of 64 (default 0 = auto)
//     --schedule plain|learned   (learned needs --model)
//     --model FILE          learned-schedule weights (5378 little-endian float32)
//     --sweeps N            sweep rounds before the queue phase (default: schedule)
//     --repeat N            reconstruct N times from the input marker, print each time
//     --legacy              call MorphologicalReconstruction (square uint8, 4-connected)
//     --stats               print the run's counters
//     --workers-per-cu N    persistent workgroups per CU in the queue phase (0 = occupancy)
//
// Like the reference, the reconstruction alone is timed (the marker is
// clipped inside the call, the reference clips before its timer).  Exit code
// 0 on success, 1 on bad usage or I/O, 2 on a library error.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "iwpp.h"
#include "iwpp_io.h"

namespace {

int usage() {
    fprintf(stderr,
            "usage: iwpp_recon MARKER.pgm MASK.pgm OUT.pgm [--conn 4|8] "
            "[--queue mutex|dq|uq|auto] [--tile N] [--schedule plain|learned] [--model FILE] "
            "[--sweeps N] [--repeat N] [--legacy] [--stats] [--workers-per-cu N]\n");
    return 1;
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
}

struct HostImage {
    void* data = nullptr;
    int w = 0, h = 0, maxval = 0;
    ~HostImage() { iwpp_pgm_free(data); }
};

bool read_image(const char* path, HostImage& im) {
    const int rc = iwpp_pgm_read(path, &im.data, &im.w, &im.h, &im.maxval);
    if (rc != IWPP_IO_OK) {
        fprintf(stderr, "cannot read %s (code %d)\n", path, rc);
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) return usage();
    const char* marker_path = argv[1];
    const char* mask_path = argv[2];
    const char* out_path = argv[3];
    int conn = 4, tile = 0, queue = IWPP_QUEUE_AUTO, schedule = IWPP_SCHED_PLAIN, sweeps = -1;
    int repeat = 1, wpc = 0;
    bool legacy = false, stats = false;
    const char* model = nullptr;
    for (int i = 4; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--conn" && (v = next())) conn = atoi(v);
        else if (a == "--tile" && (v = next())) tile = atoi(v);
        else if (a == "--sweeps" && (v = next())) sweeps = atoi(v);
        else if (a == "--repeat" && (v = next())) repeat = atoi(v);
        else if (a == "--workers-per-cu" && (v = next())) wpc = atoi(v);
        else if (a == "--model" && (v = next())) model = v;
        else if (a == "--queue" && (v = next())) {
            const std::string q = v;
            queue = q == "mutex" ? IWPP_QUEUE_MUTEX : q == "dq" ? IWPP_QUEUE_DQ
                  : q == "uq" ? IWPP_QUEUE_UQ : q == "auto" ? IWPP_QUEUE_AUTO : -1;
            if (queue < 0) return usage();
        } else if (a == "--schedule" && (v = next())) {
            const std::string q = v;
            schedule = q == "plain" ? IWPP_SCHED_PLAIN : q == "learned" ? IWPP_SCHED_LEARNED : -1;
            if (schedule < 0) return usage();
        } else if (a == "--legacy") legacy = true;
        else if (a == "--stats") stats = true;
        else return usage();
    }
    if ((conn != 4 && conn != 8) || repeat < 1) return usage();

    HostImage mk, ms;
    if (!read_image(marker_path, mk) || !read_image(mask_path, ms)) return 1;
    if (mk.w != ms.w || mk.h != ms.h) {
        fprintf(stderr, "marker %dx%d and mask %dx%d differ in size\n", mk.w, mk.h, ms.w, ms.h);
        return 1;
    }
    const int dtype = (mk.maxval > 255 || ms.maxval > 255) ? IWPP_U16 : IWPP_U8;
    const int w = mk.w, h = mk.h;
    const size_t row = (size_t)w * dtype;
    // Both planes in the wider pixel type.
    std::vector<uint8_t> hm(row * h), hk(row * h);
    auto widen = [&](const HostImage& src, std::vector<uint8_t>& dst) {
        const size_t n = (size_t)w * h;
        const bool src16 = src.maxval > 255;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t v = src16 ? ((const uint16_t*)src.data)[i] : ((const uint8_t*)src.data)[i];
            if (dtype == IWPP_U16) ((uint16_t*)dst.data())[i] = (uint16_t)v;
            else dst[i] = (uint8_t)v;
        }
    };
    widen(mk, hm);
    widen(ms, hk);
    if (legacy && (dtype != IWPP_U8 || w != h || conn != 4)) {
        fprintf(stderr, "--legacy needs a square 8-bit image and 4-connectivity\n");
        return 1;
    }
    if (model) {
        FILE* f = fopen(model, "rb");
        std::vector<float> wts(IWPP_MLP_PARAMS);
        const bool ok = f && fread(wts.data(), sizeof(float), wts.size(), f) == wts.size();
        if (f) fclose(f);
        if (!ok) {
            fprintf(stderr, "cannot read model %s\n", model);
            return 1;
        }
        if (iwpp_sched_set_model(wts.data(), (int)wts.size()) != IWPP_OK) {
            fprintf(stderr, "%s\n", iwpp_last_error());
            return 2;
        }
    }

    void *d_marker = nullptr, *d_mask = nullptr, *d_input = nullptr;
    size_t pm = 0, pk = 0, pi = 0;
    if (!hip_ok(hipMallocPitch(&d_marker, &pm, row, h), "hipMallocPitch") ||
        !hip_ok(hipMallocPitch(&d_mask, &pk, row, h), "hipMallocPitch") ||
        !hip_ok(hipMallocPitch(&d_input, &pi, row, h), "hipMallocPitch") ||
        !hip_ok(hipMemcpy2D(d_input, pi, hm.data(), row, row, h, hipMemcpyHostToDevice), "upload") ||
        !hip_ok(hipMemcpy2D(d_mask, pk, hk.data(), row, row, h, hipMemcpyHostToDevice), "upload"))
        return 2;
    iwpp_workspace* ws = iwpp_workspace_create((size_t)w * h, dtype);
    iwpp_options opt;
    iwpp_default_options(&opt);
    opt.sweep_rounds = sweeps;
    opt.workers_per_cu = wpc;
    int rc = IWPP_OK;
    for (int it = 0; it < repeat && rc == IWPP_OK; ++it) {
        if (!hip_ok(hipMemcpy2D(d_marker, pm, d_input, pi, row, h, hipMemcpyDeviceToDevice), "copy") ||
            !hip_ok(hipDeviceSynchronize(), "sync"))
            return 2;
        iwpp_stats st{};
        const auto t0 = std::chrono::steady_clock::now();
        if (legacy) {
            MorphologicalReconstruction((unsigned char*)d_marker, (unsigned char*)d_mask, pm, pk, w);
            if (!hip_ok(hipDeviceSynchronize(), "MorphologicalReconstruction")) return 2;
        } else {
            rc = iwpp_reconstruct_ex(d_marker, d_mask, pm, pk, w, h, dtype, conn, tile, queue,
                                     schedule, &opt, &st, ws, nullptr);
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (rc != IWPP_OK) break;
        printf("%.3f\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
        if (stats && !legacy)
            printf("visits %lld seeded %lld sweep_rounds %d queue %d tile %d schedule %d "
                   "ms_sweep %.3f ms_seed %.3f ms_queue %.3f\n",
                   st.tile_visits, st.seeded, st.sweep_rounds, st.queue, st.tile, st.schedule,
                   st.ms_sweep, st.ms_seed, st.ms_queue);
    }
    if (rc != IWPP_OK) {
        fprintf(stderr, "iwpp_reconstruct_ex: %s\n", iwpp_last_error());
        return 2;
    }
    if (!hip_ok(hipMemcpy2D(hm.data(), row, d_marker, pm, row, h, hipMemcpyDeviceToHost), "download"))
        return 2;
    iwpp_workspace_destroy(ws);
    (void)hipFree(d_marker);
    (void)hipFree(d_mask);
    (void)hipFree(d_input);
    const int out_max = dtype == IWPP_U16 ? std::max(std::max(mk.maxval, ms.maxval), 256) : 255;
    const int wrc = iwpp_pgm_write(out_path, hm.data(), row, w, h, out_max);
    if (wrc != IWPP_IO_OK) {
        fprintf(stderr, "cannot write %s (code %d)\n", out_path, wrc);
        return 1;
    }
    return 0;
}
