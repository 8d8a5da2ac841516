// Plain heuristic schedule, from measurements on MI355X with raster sweep
// rounds (two passes per round, each followed by its fix-up pass): whole
// reconstructions with k = 0 and 1 uniform rounds (ms), mutex queue, images
// of the four single-image generators (uint8 h-dome, 0.1% seeds, seeds with
// disks; uint16 h-dome).  uint8: 3072^2 0.249 0.293; 4096^2 0.298 0.340;
// 6144^2 in total 1.361 1.566; 8192^2 (three generators and four more seed
// densities and blurs) 6.130 6.324, k = 0 best on six of nine (one tie);
// 10240^2 2.267 2.291; 11264^2 2.714 2.421 (k = 1 best on each); 32768^2
// seeds 5.83 2.19.  uint16 h-dome: 3072^2 0.228 0.237; 4096^2 0.442 0.367;
// 5120^2 0.427 0.445; 6144^2 0.387 0.364; 8192^2 0.540 0.397; 10240^2 0.910
// 0.645; 12288^2 1.162 0.731; 16384^2 1.873 0.901.  So one image takes no
// sweep round while its longer side is below 10240 pixels (uint8) or 6144
// (uint16) -- its queue phase is short and a round costs more than it saves
// -- and one round from there on; the longer side decides, since a long thin
// image would otherwise leave the queue one long chain of megapixels.  A
// batch takes one round (its
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
    (void)conn;
    (void)schedule;
    if (!out || w <= 0 || h <= 0 || nimages <= 0) return -1;
    const int side = std::min(w, h), longer = std::max(w, h);
    out->tile = tile > 0 ? tile : (side >= 8192 ? 2048 : 1024);
    out->queue = queue == IWPP_QUEUE_AUTO ? IWPP_QUEUE_MUTEX : queue;
    const int no_round_below = dtype == IWPP_U16 ? 6144 : 10240;
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds
                                          : (nimages == 1 && longer < no_round_below ? 0 : 1);
    return 0;
}

}  // namespace iwpp
