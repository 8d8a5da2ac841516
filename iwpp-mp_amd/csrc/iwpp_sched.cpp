// Plain heuristic schedule, from measurements on MI355X: the mutex variant was
// fastest or tied on 4096^2 u8 seeds and 4096^2 u16 h-dome images.  Whole
// reconstructions with k = 0, 1, 2, 3 uniform sweep rounds (ms): 1024^2 h-dome
// 0.178 0.176 0.197 0.216; 4096^2 seeds 0.353 0.351 0.353 0.387; 6144^2 seeds
// 0.496 0.518 0.494 0.544; 8192^2 seeds 0.698 0.690 0.596 0.668; 8192^2 h-dome
// 0.707 0.684 0.642 0.657; 16384^2 u16 h-dome 1.97 1.51 1.97 2.68; 32768^2
// seeds 6.24 3.09 4.07 5.48.  Here: 2 rounds for short sides of 6144..16383
// pixels, 1 otherwise.  The tile is 1024 (2048 when the short side is >= 8192).
#include "iwpp_sched.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "iwpp.h"

namespace iwpp {

namespace {
std::mutex g_tune_mu;
std::map<TuneKey, TuneValue> g_tune;
int ilog2(long long v) {
    int k = 0;
    while (v > 1) {
        v >>= 1;
        ++k;
    }
    return k;
}
}  // namespace

bool TuneKey::operator<(const TuneKey& o) const {
    return std::tie(dtype, conn, batch, log2_px, log2_side, seed_class) <
           std::tie(o.dtype, o.conn, o.batch, o.log2_px, o.log2_side, o.seed_class);
}
int seed_class(unsigned long long nonzero, unsigned long long sampled) {
    if (sampled == 0) return 0;
    if (nonzero * 100 < sampled) return 0;
    return nonzero * 2 < sampled ? 1 : 2;
}
TuneKey tune_key(int dtype, int conn, int n, long long px, int max_side, int sc) {
    return TuneKey{dtype, conn, n > 1 ? 1 : 0, ilog2(px), ilog2(max_side), sc};
}
void tune_store(const TuneKey& k, const TuneValue& v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[k] = v;
}
bool tune_lookup(const TuneKey& k, TuneValue* v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(k);
    if (it == g_tune.end()) return false;
    *v = it->second;
    return true;
}
bool tune_any() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return !g_tune.empty();
}
void tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

int plan_schedule(int w, int h, int dtype, int conn, int tile, int queue, int schedule,
                  int sweep_rounds, SchedulePlan* out) {
    (void)dtype;
    (void)conn;
    (void)schedule;
    if (!out || w <= 0 || h <= 0) return -1;
    const int side = std::min(w, h);
    out->tile = tile > 0 ? tile : (side >= 8192 ? 2048 : 1024);
    out->queue = queue == IWPP_QUEUE_AUTO ? IWPP_QUEUE_MUTEX : queue;
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds : (side >= 6144 && side < 16384 ? 2 : 1);
    return 0;
}

}  // namespace iwpp
