// Plain heuristic schedule.  Measured on MI355X (4096^2 u8 seeds, 4096^2 u16
// h-dome): the mutex variant was fastest or tied, and three sweep rounds
// before the queue phase minimised the total for images of 2048 px and up;
// smaller images are dominated by fixed costs and take one round.  The tile
// is 1024 (2048 when the short side is at least 8192).
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
    out->queue = queue == IWPP_QUEUE_AUTO ? IWPP_QUEUE_MUTEX : queue;
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds : (side >= 2048 ? 3 : 1);
    return 0;
}

}  // namespace iwpp
