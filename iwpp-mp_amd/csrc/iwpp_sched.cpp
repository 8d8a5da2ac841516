// Plain heuristic schedule, from measurements on MI355X with raster sweep
// rounds (two passes per round): whole reconstructions with k = 0, 1, 2
// uniform rounds (ms), mutex queue: 1024^2 h-dome 0.187 0.185 0.214;
// 2048^2 h-dome 0.199 0.212 0.241; 4096^2 0.1% seeds 0.311 0.339 0.359;
// 6144^2 seeds 0.472 0.528 0.494; 8192^2 (13 images of both generators,
// seed densities 30-1000 ppm, blur 8-32, u8 and u16) in total 8.61 8.28 8.38
// with the best k per image 0, 1 or 2; 12288^2 seeds 1.153 1.081 1.13;
// 16384^2 u16 h-dome 1.785 1.013 1.329; 32768^2 seeds 6.31 2.50 3.37.  So:
// one image whose sides are both below 8192 pixels takes no sweep round (its
// queue phase is short: a round costs more than it saves), any other one
// round; the longer side decides, since a long thin image would otherwise
// leave the queue one long chain of megapixels.  A batch takes one round (its
// single queue phase serves every image: 256 images of 256..4096 pixels a
// side, 1 round 2.03 ms against more without).  The tile is 1024 (2048 when
// the short side is >= 8192).
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

int plan_schedule(int w, int h, int nimages, int dtype, int conn, int tile, int queue,
                  int schedule, int sweep_rounds, SchedulePlan* out) {
    (void)dtype;
    (void)conn;
    (void)schedule;
    if (!out || w <= 0 || h <= 0 || nimages <= 0) return -1;
    const int side = std::min(w, h), longer = std::max(w, h);
    out->tile = tile > 0 ? tile : (side >= 8192 ? 2048 : 1024);
    out->queue = queue == IWPP_QUEUE_AUTO ? IWPP_QUEUE_MUTEX : queue;
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds
                                          : (nimages == 1 && longer < 8192 ? 0 : 1);
    return 0;
}

}  // namespace iwpp
