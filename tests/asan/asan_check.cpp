// Host-side memory and undefined-behaviour check (the counterpart of the
// reference's compute-sanitizer targets, IWPP MP/Makefile:51-61, for the code
// that runs on the CPU): the PGM reader/writer, the synthetic generators, the
// schedule planner and tuning cache, and the CPU oracle, built with
// -fsanitize=address,undefined and driven over ragged shapes, both pixel types,
// malformed files and random cases.  Any sanitizer report fails the run.
// Device code cannot be checked this way (GPU address sanitizer builds are
// not available on the test machines).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "iwpp.h"
#include "iwpp_io.h"
#include "iwpp_sched.h"
#include "iwpp_synth.h"
extern "C" {
#include "iwpp_oracle.h"
#include "queue_model.h"
}

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() {  // xorshift64*
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (uint32_t)((rng_state * 2685821657736338717ull) >> 32);
}

static void pgm_checks(const std::string& dir) {
    const int shapes[][2] = {{1, 1}, {3, 7}, {64, 33}, {257, 5}};
    for (auto& s : shapes)
        for (int maxval : {255, 7, 4095, 65535}) {
            const int w = s[0], h = s[1];
            const size_t bpp = maxval > 255 ? 2 : 1, pitch = w * bpp + 3 * bpp;  // padded pitch
            std::vector<uint8_t> img(pitch * h);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const uint32_t v = rnd() % (maxval + 1);
                    if (bpp == 2) memcpy(&img[y * pitch + 2 * x], &v, 2);
                    else img[y * pitch + x] = (uint8_t)v;
                }
            const std::string p = dir + "/t.pgm";
            CHECK(iwpp_pgm_write(p.c_str(), img.data(), pitch, w, h, maxval) == IWPP_IO_OK);
            void* data = nullptr;
            int W, H, M;
            CHECK(iwpp_pgm_read(p.c_str(), &data, &W, &H, &M) == IWPP_IO_OK);
            CHECK(W == w && H == h && M == maxval);
            for (int y = 0; y < h && data; ++y)
                CHECK(memcmp((uint8_t*)data + y * w * bpp, &img[y * pitch], w * bpp) == 0);
            iwpp_pgm_free(data);
        }
    const char* bad[] = {"", "P5", "P5\n3", "P5\n3 2\n", "P5\n3 2\n255\n\x01", "P2\n2 2\n9\n1 2 3",
                         "P5\n-1 2\n255\n", "P5\n99999999999 2\n255\n", "P7\n1 1\n1\n\x00",
                         "P2\n1 1\n9\n# only a comment"};
    for (const char* b : bad) {
        const std::string p = dir + "/bad.pgm";
        FILE* f = fopen(p.c_str(), "wb");
        fwrite(b, 1, strlen(b), f);
        fclose(f);
        void* data = nullptr;
        int W, H, M;
        CHECK(iwpp_pgm_read(p.c_str(), &data, &W, &H, &M) != IWPP_IO_OK);
        CHECK(data == nullptr);
    }
}

// Brute-force fixed point J <- min(dilate(J), I), absent pixels outside.
template <class T>
static std::vector<T> brute(std::vector<T> J, const std::vector<T>& I, int w, int h, int conn) {
    for (size_t i = 0; i < J.size(); ++i) J[i] = std::min(J[i], I[i]);
    bool ch = true;
    while (ch) {
        ch = false;
        std::vector<T> N = J;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                T m = J[y * w + x];
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (conn == 4 && dx && dy) continue;
                        const int yy = y + dy, xx = x + dx;
                        if (yy >= 0 && yy < h && xx >= 0 && xx < w) m = std::max(m, J[yy * w + xx]);
                    }
                m = std::min(m, I[y * w + x]);
                if (m != N[y * w + x]) {
                    N[y * w + x] = m;
                    ch = true;
                }
            }
        J.swap(N);
    }
    return J;
}

template <class T>
static void oracle_checks(int dtype) {
    const int shapes[][2] = {{1, 1}, {1, 37}, {37, 1}, {13, 17}, {64, 65}, {100, 3}};
    for (auto& s : shapes)
        for (int conn : {4, 8}) {
            const int w = s[0], h = s[1];
            std::vector<T> I(w * h), J(w * h);
            const uint32_t top = dtype == IWPP_U8 ? 255 : 4095;
            for (auto& v : I) v = (T)(rnd() % (top + 1));
            for (auto& v : J) v = (rnd() % 17 == 0) ? (T)(rnd() % (top + 1)) : 0;
            std::vector<T> want = brute(J, I, w, h, conn);
            CHECK(oracle_reconstruct(J.data(), I.data(), w * sizeof(T), w * sizeof(T), w, h,
                                     dtype, conn) == 0);
            CHECK(J == want);
        }
}

static void synth_checks() {
    for (int dtype : {IWPP_U8, IWPP_U16}) {
        const int w = 97, h = 61, maxval = dtype == IWPP_U8 ? 255 : 4095;
        const size_t pitch = w * dtype + 6;
        std::vector<uint8_t> mask(pitch * h), marker(pitch * h);
        CHECK(iwpp_synth_mask(mask.data(), pitch, w, h, dtype, maxval, 4.0, 3, 5, 20, 7, 2) == 0);
        CHECK(iwpp_synth_marker_hdome(marker.data(), mask.data(), pitch, pitch, w, h, dtype, 30) == 0);
        CHECK(iwpp_synth_marker_seeds(marker.data(), mask.data(), pitch, pitch, w, h, dtype, 50, 9) == 0);
    }
    std::vector<int> ws(40), hs(40), ks(40), hv(40);
    CHECK(iwpp_synth_batch_shapes(40, 3, ws.data(), hs.data(), ks.data(), hv.data()) == 0);
    std::vector<uint8_t> a(33 * 20), b(33 * 20);
    CHECK(iwpp_synth_random_case(a.data(), b.data(), 33, 33, 33, 20, 5) == 0);
    // misaligned uint16 rows are refused, not written
    std::vector<uint16_t> c(64);
    CHECK(iwpp_synth_mask((uint8_t*)c.data() + 1, 16, 4, 4, 2, 4095, 1.0, 0, 0, 0, 1, 1) < 0);
    CHECK(iwpp_pgm_write("/dev/null", (uint8_t*)c.data() + 1, 16, 4, 4, 4095) == IWPP_IO_ERR_ARG);
}

static void sched_checks() {
    iwpp::SchedulePlan p;
    for (int w : {1, 100, 4096, 6144, 8192, 10240, 32768})
        for (int q : {IWPP_QUEUE_MUTEX, IWPP_QUEUE_DQ, IWPP_QUEUE_UQ, IWPP_QUEUE_AUTO}) {
            for (int n : {1, 7}) {
                CHECK(iwpp::plan_schedule(w, w, n, IWPP_U8, 8, 0, q, IWPP_SCHED_PLAIN, -1, &p) == 0);
                CHECK(p.tile % 64 == 0 && p.sweep_rounds >= 0);
                CHECK(p.sweep_rounds == (n == 1 && w < 10240 ? 0 : 1));
                CHECK(iwpp::plan_schedule(w, w, n, IWPP_U16, 4, 0, q, IWPP_SCHED_PLAIN, -1, &p) == 0);
                CHECK(p.sweep_rounds == (n == 1 && w < 6144 ? 0 : 1));
            }
        }
    CHECK(iwpp::plan_schedule(1, 100000, 1, IWPP_U8, 8, 0, 0, IWPP_SCHED_PLAIN, -1, &p) == 0 &&
          p.sweep_rounds == 1);  // a long thin image sweeps
    CHECK(iwpp::plan_schedule(64, 64, 0, IWPP_U8, 8, 0, 0, IWPP_SCHED_PLAIN, -1, &p) < 0);
    iwpp::tune_store(iwpp::tune_key(1, 8, 1, 1 << 20, 1024, 0), iwpp::TuneValue{2, 3, 256});
    iwpp::TuneValue tv{};
    CHECK(iwpp::tune_lookup(iwpp::tune_key(1, 8, 1, 1 << 20, 1024, 0), &tv) && tv.queue == 2 &&
          tv.sweep_rounds == 3 && tv.col_block_rows == 256);
    CHECK(!iwpp::tune_lookup(iwpp::tune_key(1, 8, 1, 1 << 20, 1024, 2), &tv));  // other seed class
    CHECK(iwpp::seed_class(0, 100) == 0 && iwpp::seed_class(5, 100) == 1 &&
          iwpp::seed_class(60, 100) == 2 && iwpp::seed_class(0, 0) == 0);
    iwpp::tune_clear();
    CHECK(!iwpp::tune_lookup(iwpp::tune_key(1, 8, 1, 1 << 20, 1024, 0), &tv));
}

// The CPU queue-phase model: the reference's queue KATs, and every variant
// over ragged shapes, both pixel types, small queue capacities (overflow and
// restart) and capped local sweeps, against the oracle.
template <class T>
static void queue_model_checks(int dtype) {
    CHECK(qm_dq_fill_swap_kat(1) == 0);
    CHECK(qm_dq_fill_swap_kat(64) == 0);
    CHECK(qm_uq_unique_kat(16, 3, 7) == 0);
    const int shapes[][2] = {{1, 1}, {37, 1}, {1, 53}, {45, 29}, {64, 70}};
    for (auto& s : shapes)
        for (int conn : {4, 8})
            for (int variant : {QM_MUTEX, QM_DQ, QM_UQ}) {
                const int w = s[0], h = s[1];
                std::vector<T> I(w * h), J(w * h);
                for (auto& v : I) v = (T)(rnd() % (dtype == 1 ? 256 : 4096));
                for (size_t k = 0; k < J.size(); ++k) J[k] = rnd() % 50 == 0 ? I[k] : 0;
                std::vector<T> want = J, got = J;
                CHECK(oracle_reconstruct(want.data(), I.data(), w * sizeof(T), w * sizeof(T), w, h,
                                         dtype, conn) == 0);
                qm_stats st;
                CHECK(qm_reconstruct(got.data(), I.data(), w * sizeof(T), w * sizeof(T), w, h,
                                     dtype, conn, variant, 8, 5, rnd(),
                                     variant == QM_MUTEX ? 0 : 3, 4, QM_FAULT_NONE, &st) == 0);
                CHECK(got == want);
                CHECK(st.state_left == 0 && st.overlap == 0);
            }
    std::vector<T> z(4);
    qm_stats st;
    CHECK(qm_reconstruct(z.data(), z.data(), 2 * sizeof(T), 2 * sizeof(T), 2, 2, dtype, 6, 0, 8,
                         1, 1, 0, 0, 0, &st) < 0);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    pgm_checks(dir);
    synth_checks();
    sched_checks();
    oracle_checks<uint8_t>(IWPP_U8);
    oracle_checks<uint16_t>(IWPP_U16);
    queue_model_checks<uint8_t>(IWPP_U8);
    queue_model_checks<uint16_t>(IWPP_U16);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("asan_check: all checks passed\n");
    return 0;
}
