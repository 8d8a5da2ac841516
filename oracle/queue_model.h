/*
 * CPU model of the three queue disciplines of the IWPP queue phase.
 *
 * TEST INFRASTRUCTURE ONLY (like the rest of oracle/): the checker for what
 * pixel comparisons cannot show -- how work moves through the queues.  The
 * product path never links or calls it.
 *
 * It restates, on the CPU and with simulated concurrency, the semantics of
 *   - the mutex queue: one global queue drained by persistent workers, each
 *     megapixel queued at most once and never visited by two workers at once
 *     (this build's ring with PENDING/RUNNING state bits, which replaces the
 *     reference's spin-locked stack, IWPP MP/src/globalQueue.cu:25-154);
 *   - the double queue: read and write queues swapped between rounds,
 *     duplicates allowed, a megapixel that rose re-queues itself
 *     (IWPP MP DQ/src/globalQueue.cu:14-157, IWPP MP DQ/src/
 *     MorphologicalReconstruction.cu:1242-1323);
 *   - the unique queue: the double queue with at most one entry per megapixel
 *     per round (IWPP MP UQ/src/globalQueue.cu:159-168, IWPP MP UQ/src/
 *     MorphologicalReconstruction.cu:1152-1293; here the flag is set at push
 *     and cleared at pop, as the GPU kernel does);
 * including the overflow restart from the partial marker (the reference's
 * memLeak loop).  A visit is the GPU's: snapshot of the megapixel and its
 * one-pixel halo, local sweeps, write-back (exclusive, or max-merge for the
 * double queue), notification of the neighbours whose halo pixels can be
 * raised by a border pixel that rose.  Workers advance one step at a time in
 * an order drawn from a seeded generator, so the snapshots of concurrent
 * visits are stale in the ways the GPU's can be.
 */
#ifndef IWPP_QUEUE_MODEL_H
#define IWPP_QUEUE_MODEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { QM_MUTEX = 0, QM_DQ = 1, QM_UQ = 2 };
/* Protocol mistakes a test can plant to show that the model's checks see them:
 * the mutex variant forgetting a megapixel notified while it ran, the unique
 * queue never clearing a megapixel's queued flag when it is taken. */
enum { QM_FAULT_NONE = 0, QM_FAULT_MUTEX_DROP_RENOTIFY = 1, QM_FAULT_UQ_KEEP_FLAG = 2 };

typedef struct {
    long long visits;          /* megapixel visits */
    long long seeded;          /* megapixels put in the queue by the seeding passes */
    long long rounds;          /* DQ / UQ rounds */
    long long restarts;        /* queue overflows that re-seeded from the partial marker */
    long long overlap;         /* mutex: visits that overlapped a visit of the same megapixel */
    long long repeat_in_round; /* DQ / UQ: visits of a megapixel already visited that round */
    long long state_left;      /* state words / queued flags not clear at the end */
    long long steps;           /* scheduler steps */
} qm_stats;

/* Reconstructs marker in place under mask through the given queue variant.
 * mp: megapixel edge (pixels); workers: simulated concurrent workers;
 * seed: the interleaving; cap: queue capacity in entries (0 = ample; small
 * values force overflows and restarts); max_sweeps: local sweeps per visit
 * before a megapixel re-queues itself (0 = until stable, else at least 4:
 * a visit certifies convergence with a quiet sweep in each direction);
 * fault: QM_FAULT_* (tests only; 0 = the protocol as specified).
 * Returns 0, or a negative value on bad arguments / allocation failure. */
int qm_reconstruct(void* marker, const void* mask, size_t pitch_j, size_t pitch_i, int width,
                   int height, int dtype, int conn, int variant, int mp, int workers,
                   uint64_t seed, long cap, int max_sweeps, int fault, qm_stats* stats);

/* The reference's double-queue known-answer test (IWPP MP DQ/src/
 * globalQueue.cu:218-253) on the model's queue operations: every coordinate
 * of a g x g grid appended to the read queue four times, swap, once, swap,
 * twice; returns how many coordinates do not appear exactly twice in the
 * read queue (0 = pass), or a negative value on failure to allocate. */
long qm_dq_fill_swap_kat(int g);

/* The unique queue's flag table check (IWPP MP UQ/src/
 * MorphologicalReconstruction.cu:1152-1293, 1312-1322): a write queue with
 * every coordinate of a g x g grid `dup` times, in a seeded order, has its
 * unique entries extracted with the flags and the flags cleared; returns
 * the number of flags left nonzero plus the number of coordinates not
 * exactly once in the new read queue (0 = pass). */
long qm_uq_unique_kat(int g, int dup, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
