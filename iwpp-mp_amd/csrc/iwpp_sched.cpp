// Plain heuristic schedule, from measurements on MI355X: the mutex variant was
// fastest or tied on 4096^2 u8 seeds and 4096^2 u16 h-dome images; the best
// number of sweep rounds before the queue phase was 3 for 4096^2 seeds, 2 for
// 32768^2 seeds and 1 for 16384^2 h-dome -- data dependent, which is what the
// learned schedule decides per tile.  Here: 1 round below 2048 px, 3 up to
// 8191, 2 from 8192.  The tile is 1024 (2048 when the short side is >= 8192).
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
    out->sweep_rounds = sweep_rounds >= 0 ? sweep_rounds : (side >= 8192 ? 2 : side >= 2048 ? 3 : 1);
    return 0;
}

}  // namespace iwpp
