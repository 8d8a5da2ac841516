// Plain heuristic schedule, from measurements on MI355X: the mutex variant was
// fastest or tied on 4096^2 u8 seeds and 4096^2 u16 h-dome images; the best
// number of sweep rounds before the queue phase was 3 for 4096^2 seeds, 2 for
// 32768^2 seeds and 1 for 16384^2 h-dome -- data dependent, which is what the
// learned schedule decides per tile.  Here: 1 round below 2048 px, 3 up to
// 8191, 2 from 8192.  The tile is 1024 (2048 when the short side is >= 8192).
#include "iwpp_sched.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "iwpp.h"

namespace iwpp {

namespace {
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune;
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
    return std::tie(dtype, conn, batch, log2_px, log2_side) <
           std::tie(o.dtype, o.conn, o.batch, o.log2_px, o.log2_side);
}
TuneKey tune_key(int dtype, int conn, int n, long long px, int max_side) {
    return TuneKey{dtype, conn, n > 1 ? 1 : 0, ilog2(px), ilog2(max_side)};
}
void tune_store(const TuneKey& k, int queue, int sweep_rounds) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[k] = {queue, sweep_rounds};
}
bool tune_lookup(const TuneKey& k, int* queue, int* sweep_rounds) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(k);
    if (it == g_tune.end()) return false;
    *queue = it->second.first;
    *sweep_rounds = it->second.second;
    return true;
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
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds : (side >= 8192 ? 2 : side >= 2048 ? 3 : 1);
    return 0;
}

}  // namespace iwpp
