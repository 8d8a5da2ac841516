// Plain heuristic schedule.  The queue variant defaults to UQ, the sweep
// phase to one round, the tile to 1024 (2048 for images whose short side is
// at least 8192).
#include "iwpp_sched.h"

#include <algorithm>

#include "iwpp.h"

namespace iwpp {

int plan_schedule(int w, int h, int dtype, int conn, int tile, int queue, int schedule,
                  int sweep_rounds, SchedulePlan* out) {
    (void)dtype;
    (void)conn;
    (void)schedule;
    if (!out || w <= 0 || h <= 0) return -1;
    const int side = std::min(w, h);
    out->tile = tile > 0 ? tile : (side >= 8192 ? 2048 : 1024);
    out->queue = queue == IWPP_QUEUE_AUTO ? IWPP_QUEUE_UQ : queue;
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds : 1;
    return 0;
}

}  // namespace iwpp
