// Schedule planning: tile size, queue variant and sweep rounds before the
// queue phase.  Host only, no device code.
#pragma once

namespace iwpp {

struct SchedulePlan {
    int tile;          // scheduling tile edge in pixels (multiple of 64)
    int queue;         // Q_MUTEX / Q_DQ / Q_UQ
    int sweep_rounds;  // sweep rounds before the queue phase
};

// Fills *out for one image of w x h (nimages = 1) or a batch of nimages
// whose widest image is w and tallest h.  tile = 0 and queue =
// IWPP_QUEUE_AUTO ask for the planner's choice; sweep_rounds < 0 asks for the
// schedule's default.  Returns 0, or a negative value for an impossible
// request.
int plan_schedule(int w, int h, int nimages, int dtype, int conn, int tile, int queue,
                  int schedule, int sweep_rounds, SchedulePlan* out);

// Autotuner cache: the measured best (queue, sweep rounds, column block
// height) for a workload key (dtype, connectivity, single image or batch,
// log2 of the pixel count, log2 of the widest side, seed class).  The
// reconstruction uses it for queue = AUTO / sweep_rounds < 0 /
// col_block_rows = 0.
struct TuneKey {
    int dtype, conn, batch, log2_px, log2_side, seed_class;
    bool operator<(const TuneKey& o) const;
};
struct TuneValue {
    int queue, sweep_rounds, col_block_rows;
};
// Seed class of a marker from the fraction f of (sampled) pixels where
// min(marker, mask) > 0: 0 = sparse seeds (f < 1%), 1 = partial (f < 50%),
// 2 = dense (h-dome-like markers).
int seed_class(unsigned long long nonzero, unsigned long long sampled);
TuneKey tune_key(int dtype, int conn, int n, long long px, int max_side, int seed_class);
void tune_store(const TuneKey& k, const TuneValue& v);
bool tune_lookup(const TuneKey& k, TuneValue* v);
bool tune_any();
void tune_clear();

}  // namespace iwpp
