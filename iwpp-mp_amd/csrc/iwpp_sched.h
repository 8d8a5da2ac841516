// Schedule planning: tile size, queue variant and sweep rounds before the
// queue phase.  Host only, no device code.
#pragma once

namespace iwpp {

struct SchedulePlan {
    int tile;          // scheduling tile edge in pixels (multiple of 64)
    int queue;         // Q_MUTEX / Q_DQ / Q_UQ
    int sweep_rounds;  // rounds of (row pass, column pass down, column pass up)
};

// Fills *out for one image.  tile = 0 and queue = IWPP_QUEUE_AUTO ask for the
// planner's choice; sweep_rounds < 0 asks for the schedule's default.
// Returns 0, or a negative value for an impossible request.
int plan_schedule(int w, int h, int dtype, int conn, int tile, int queue, int schedule,
                  int sweep_rounds, SchedulePlan* out);

}  // namespace iwpp
