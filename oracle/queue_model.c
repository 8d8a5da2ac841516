/*
 * CPU model of the IWPP queue phase's three queue disciplines (declarations
 * and the reference citations are in queue_model.h).
 * TEST INFRASTRUCTURE ONLY: a checker, never part of the product path.
 *
 * Pixels are held as uint16 (uint8 images are widened); the update rule and
 * zero halos are the reference's (IWPP MP/src/MorphologicalReconstruction.cu
 * :42-47, :252-312).  A visit: snapshot the megapixel and its halo, sweep
 * the snapshot to a fixed point (or max_sweeps sweeps), write it back,
 * notify neighbours.  Scheduling: `workers` state machines advanced one step
 * at a time in a seeded random order; the steps of a visit are
 *   take -> fetch (snapshot) -> store (write-back) -> notify/release,
 * so other workers' stores land between a visit's snapshot and its
 * write-back exactly as they can on the GPU.
 */
#include "queue_model.h"

#include <stdlib.h>
#include <string.h>

#define MAXV(a, b) ((a) > (b) ? (a) : (b))
#define MINV(a, b) ((a) < (b) ? (a) : (b))

static uint64_t rng_next(uint64_t* s) {  /* splitmix64 */
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    int w, h, conn, mp, nmx, nmy, nmp, max_sweeps;
    uint16_t *J, *I;
} img_t;

/* ------------------------------------------------------------- a visit -- */
enum { W_IDLE = 0, W_TAKEN, W_FETCHED, W_STORED };

typedef struct {
    int state, m;
    int tw;                 /* tile width with halo = mp + 2 */
    uint16_t *j0, *j, *i;   /* snapshot (fetched), working copy, mask; tw * tw */
    int conv, rose_mem;     /* converged; DQ: some pixel in memory rose */
    uint32_t nb;            /* neighbour bits to notify (bit k: k-th of 9, 4 = self) */
} worker_t;

static int in_img(const img_t* im, int x, int y) { return x >= 0 && y >= 0 && x < im->w && y < im->h; }

static void fetch(const img_t* im, worker_t* wk) {
    const int mp = im->mp, tw = wk->tw;
    const int x0 = (wk->m % im->nmx) * mp, y0 = (wk->m / im->nmx) * mp;
    for (int r = 0; r < tw; ++r)
        for (int c = 0; c < tw; ++c) {
            const int x = x0 + c - 1, y = y0 + r - 1;
            const int ok = in_img(im, x, y);
            wk->j0[r * tw + c] = ok ? im->J[(size_t)y * im->w + x] : 0;
            wk->i[r * tw + c] = ok ? im->I[(size_t)y * im->w + x] : 0;
        }
    memcpy(wk->j, wk->j0, sizeof(uint16_t) * tw * tw);
}

/* One directional sweep over the interior of the working copy; returns
 * whether a pixel rose.  dir 0: rows left->right, 1: right->left, 2: columns
 * top->bottom, 3: bottom->top; 8-connectivity takes the diagonal neighbours
 * of the previous row / column as well. */
static int sweep(const img_t* im, worker_t* wk, int dir) {
    const int tw = wk->tw, n = im->mp;
    uint16_t* J = wk->j;
    const uint16_t* I = wk->i;
    int changed = 0;
    for (int a = 1; a <= n; ++a)
        for (int s = 1; s <= n; ++s) {
            int r, c, pr, pc;
            if (dir == 0) { r = a; c = s; pr = r; pc = c - 1; }
            else if (dir == 1) { r = a; c = n + 1 - s; pr = r; pc = c + 1; }
            else if (dir == 2) { r = s; c = a; pr = r - 1; pc = c; }
            else { r = n + 1 - s; c = a; pr = r + 1; pc = c; }
            uint16_t q = J[pr * tw + pc];
            if (im->conn == 8) {
                if (dir < 2) {
                    q = MAXV(q, J[(r - 1) * tw + pc]);
                    q = MAXV(q, J[(r + 1) * tw + pc]);
                } else {
                    q = MAXV(q, J[pr * tw + c - 1]);
                    q = MAXV(q, J[pr * tw + c + 1]);
                }
            }
            const uint16_t v = MAXV(J[r * tw + c], MINV(q, I[r * tw + c]));
            if (v != J[r * tw + c]) {
                J[r * tw + c] = v;
                changed = 1;
            }
        }
    return changed;
}

static void converge(const img_t* im, worker_t* wk) {
    int quiet = 0, sweeps = 0;
    wk->conv = 0;
    for (int d = 0;; d = (d + 1) & 3) {
        if (im->max_sweeps > 0 && sweeps >= im->max_sweeps) return;
        ++sweeps;
        if (sweep(im, wk, d)) quiet = 0;
        else if (++quiet >= 4) break;  /* every direction quiet: a fixed point */
    }
    wk->conv = 1;
}

/* Write-back.  exclusive: the worker is the megapixel's only writer and
 * overwrites it with its result; otherwise (DQ) a per-pixel max-merge with
 * what memory holds now.  Then the neighbour bits: neighbour n is notified
 * when a border pixel p rose in this visit and a halo pixel q of n next to p
 * has min(J(p), I(q)) > J(q) in the snapshot. */
static void store(const img_t* im, worker_t* wk, int exclusive) {
    const int mp = im->mp, tw = wk->tw;
    const int mx = wk->m % im->nmx, my = wk->m / im->nmx, x0 = mx * mp, y0 = my * mp;
    wk->rose_mem = 0;
    for (int r = 1; r <= mp; ++r)
        for (int c = 1; c <= mp; ++c) {
            const int x = x0 + c - 1, y = y0 + r - 1;
            if (!in_img(im, x, y)) continue;
            uint16_t* g = &im->J[(size_t)y * im->w + x];
            const uint16_t v = wk->j[r * tw + c];
            if (exclusive) *g = v;
            else if (v > *g) {
                *g = v;
                wk->rose_mem = 1;
            }
        }
    wk->nb = 0;
    for (int r = 1; r <= mp; ++r)
        for (int c = 1; c <= mp; ++c) {
            if (r != 1 && r != mp && c != 1 && c != mp) continue;
            const uint16_t p = wk->j[r * tw + c];
            if (p == wk->j0[r * tw + c]) continue;  /* did not rise */
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    if (im->conn == 4 && dx && dy) continue;
                    const int rr = r + dy, cc = c + dx;
                    if (rr >= 1 && rr <= mp && cc >= 1 && cc <= mp) continue;  /* inside */
                    const int qx = x0 + cc - 1, qy = y0 + rr - 1;
                    if (!in_img(im, qx, qy)) continue;
                    const uint16_t qj = wk->j0[rr * tw + cc], qi = wk->i[rr * tw + cc];
                    if (MINV(p, qi) > qj) {
                        const int bx = rr < 1 ? 0 : rr > mp ? 2 : 1;
                        const int by = cc < 1 ? 0 : cc > mp ? 2 : 1;
                        wk->nb |= 1u << (bx * 3 + by);
                    }
                }
        }
    if (!wk->conv) wk->nb |= 1u << 4;
}

static int nb_index(const img_t* im, int m, int k) {  /* k = 3 * (dy + 1) + (dx + 1) */
    const int mx = m % im->nmx + (k % 3) - 1, my = m / im->nmx + (k / 3) - 1;
    return my * im->nmx + mx;
}

/* ------------------------------------------------------------ seeding --- */
static int unstable_mp(const img_t* im, int m) {
    const int mp = im->mp, x0 = (m % im->nmx) * mp, y0 = (m / im->nmx) * mp;
    for (int y = y0; y < MINV(y0 + mp, im->h); ++y)
        for (int x = x0; x < MINV(x0 + mp, im->w); ++x) {
            const uint16_t jp = im->J[(size_t)y * im->w + x], ip = im->I[(size_t)y * im->w + x];
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    if (im->conn == 4 && dx && dy) continue;
                    if (!in_img(im, x + dx, y + dy)) continue;
                    const uint16_t jq = im->J[(size_t)(y + dy) * im->w + x + dx];
                    if (MINV(jq, ip) > jp) return 1;
                }
        }
    return 0;
}

/* ------------------------------------------------------------ variants --- */
typedef struct {
    int variant;
    long cap;
    /* mutex */
    int* ring;
    long rhead, rtail, rcap;  /* FIFO, counts */
    uint8_t* st;               /* PENDING 1 / RUNNING 2 */
    long pending;
    int* holders;
    /* DQ / UQ */
    int *rq, *wq;
    long rsize, wsize, cursor;
    int overflow;
    int fault;                 /* planted protocol mistake (QM_FAULT_*), 0 = none */
    long long round;
    long long* last_round;
    uint8_t* flag;             /* UQ queued flag */
} queue_t;

static void ring_push(queue_t* q, int m) {
    q->ring[q->rtail % q->rcap] = m;
    q->rtail++;
}

static void wq_push(queue_t* q, int m) {
    if (q->wsize < q->cap) q->wq[q->wsize] = m;
    else q->overflow = 1;
    q->wsize++;
}

/* one worker step; returns 1 if the worker did something */
static int step(const img_t* im, queue_t* q, worker_t* wk, qm_stats* S) {
    switch (wk->state) {
        case W_IDLE:
            if (q->variant == QM_MUTEX) {
                if (q->rhead == q->rtail) return 0;
                wk->m = q->ring[q->rhead % q->rcap];
                q->rhead++;
                q->st[wk->m] = 2;  /* RUNNING (PENDING cleared) */
                if (q->holders[wk->m]++ != 0) S->overlap++;
            } else {
                if (q->cursor >= q->rsize) return 0;
                wk->m = q->rq[q->cursor++];
                if (q->variant == QM_UQ && q->fault != QM_FAULT_UQ_KEEP_FLAG) q->flag[wk->m] = 0;
                if (q->last_round[wk->m] == q->round) S->repeat_in_round++;
                q->last_round[wk->m] = q->round;
            }
            wk->state = W_TAKEN;
            return 1;
        case W_TAKEN:
            fetch(im, wk);
            wk->state = W_FETCHED;
            return 1;
        case W_FETCHED:
            converge(im, wk);
            store(im, wk, q->variant != QM_DQ);
            if (q->variant == QM_DQ && wk->rose_mem) wk->nb |= 1u << 4;  /* re-queue self */
            wk->state = W_STORED;
            S->visits++;
            return 1;
        case W_STORED:
            for (int k = 0; k < 9; ++k) {
                if (!((wk->nb >> k) & 1u)) continue;
                const int n = nb_index(im, wk->m, k);
                if (q->variant == QM_MUTEX) {
                    const uint8_t old = q->st[n];
                    q->st[n] |= 1;
                    if (old == 0) {
                        q->pending++;
                        ring_push(q, n);
                    }
                } else if (q->variant == QM_UQ) {
                    if (!q->flag[n]) {
                        q->flag[n] = 1;
                        wq_push(q, n);
                    }
                } else {
                    wq_push(q, n);
                }
            }
            if (q->variant == QM_MUTEX) {
                q->holders[wk->m]--;
                const uint8_t old = q->st[wk->m];
                q->st[wk->m] &= (uint8_t)~2u;
                if ((old & 1) && q->fault != QM_FAULT_MUTEX_DROP_RENOTIFY)
                    ring_push(q, wk->m);  /* notified while running */
                else if (old & 1) q->st[wk->m] = 0, q->pending--;  /* planted: the re-queue is lost */
                else q->pending--;
            }
            wk->state = W_IDLE;
            return 1;
    }
    return 0;
}

static int all_idle(const worker_t* ws, int n) {
    for (int i = 0; i < n; ++i)
        if (ws[i].state != W_IDLE) return 0;
    return 1;
}

/* One queue phase from the current marker; returns 1 when a queue overflowed
 * (restart), 0 when done, negative on allocation failure. */
static int phase(const img_t* im, queue_t* q, worker_t* ws, int nw, uint64_t* rs, qm_stats* S) {
    /* seeding: unstable megapixels in index order */
    long seeded = 0;
    if (q->variant == QM_MUTEX) {
        q->rhead = q->rtail = 0;
        q->pending = 0;
        memset(q->st, 0, (size_t)im->nmp);
    } else {
        q->rsize = q->wsize = q->cursor = 0;
        q->overflow = 0;
        if (q->variant == QM_UQ) memset(q->flag, 0, (size_t)im->nmp);
    }
    for (int m = 0; m < im->nmp; ++m) {
        if (!unstable_mp(im, m)) continue;
        ++seeded;
        if (q->variant == QM_MUTEX) {
            q->st[m] = 1;
            q->pending++;
            ring_push(q, m);
        } else {
            if (q->variant == QM_UQ) q->flag[m] = 1;
            wq_push(q, m);
        }
    }
    S->seeded += seeded;
    if (q->variant != QM_MUTEX) {  /* the seeds (the first cap of them) are round 0's */
        int* t = q->rq;
        q->rq = q->wq;
        q->wq = t;
        q->rsize = MINV(q->wsize, q->cap);
        q->wsize = 0;
        q->cursor = 0;
    }
    for (;;) {
        S->steps++;
        const int i = (int)(rng_next(rs) % (uint64_t)nw);
        step(im, q, &ws[i], S);
        if (q->variant == QM_MUTEX) {
            if (q->pending == 0 && all_idle(ws, nw)) return 0;
        } else if (q->cursor >= q->rsize && all_idle(ws, nw)) {  /* end of a round */
            S->rounds++;
            q->round++;
            if (q->overflow) return 1;
            if (q->wsize == 0) return 0;
            int* t = q->rq;
            q->rq = q->wq;
            q->wq = t;
            q->rsize = MINV(q->wsize, q->cap);
            q->wsize = 0;
            q->cursor = 0;
        }
    }
}

int qm_reconstruct(void* marker, const void* mask, size_t pj, size_t pi, int w, int h, int dtype,
                   int conn, int variant, int mp, int workers, uint64_t seed, long cap,
                   int max_sweeps, int fault, qm_stats* stats) {
    if (!marker || !mask || w <= 0 || h <= 0 || (dtype != 1 && dtype != 2)) return -1;
    if ((conn != 4 && conn != 8) || variant < QM_MUTEX || variant > QM_UQ || mp < 2 ||
        workers < 1 || (max_sweeps > 0 && max_sweeps < 4))
        return -1;  /* (a visit certifies convergence with four quiet sweeps) */
    if (pj < (size_t)w * dtype || pi < (size_t)w * dtype) return -1;
    qm_stats S;
    memset(&S, 0, sizeof S);
    img_t im;
    im.w = w;
    im.h = h;
    im.conn = conn;
    im.mp = mp;
    im.max_sweeps = max_sweeps;
    im.nmx = (w + mp - 1) / mp;
    im.nmy = (h + mp - 1) / mp;
    im.nmp = im.nmx * im.nmy;
    im.J = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)w * h);
    im.I = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)w * h);
    queue_t q;
    memset(&q, 0, sizeof q);
    q.variant = variant;
    q.fault = fault;
    q.cap = cap > 0 ? cap : 64L * im.nmp + 64;
    q.rcap = im.nmp + 1;
    worker_t* ws = (worker_t*)calloc((size_t)workers, sizeof(worker_t));
    int rc = -2;
    if (!im.J || !im.I || !ws) goto done;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const uint8_t* rj = (const uint8_t*)marker + (size_t)y * pj;
            const uint8_t* ri = (const uint8_t*)mask + (size_t)y * pi;
            uint16_t a, b;
            if (dtype == 1) {
                a = rj[x];
                b = ri[x];
            } else {
                memcpy(&a, rj + 2 * x, 2);
                memcpy(&b, ri + 2 * x, 2);
            }
            im.J[(size_t)y * w + x] = MINV(a, b);  /* clip */
            im.I[(size_t)y * w + x] = b;
        }
    const int tw = mp + 2;
    for (int i = 0; i < workers; ++i) {
        ws[i].tw = tw;
        ws[i].j0 = (uint16_t*)malloc(sizeof(uint16_t) * tw * tw);
        ws[i].j = (uint16_t*)malloc(sizeof(uint16_t) * tw * tw);
        ws[i].i = (uint16_t*)malloc(sizeof(uint16_t) * tw * tw);
        if (!ws[i].j0 || !ws[i].j || !ws[i].i) goto done;
    }
    if (variant == QM_MUTEX) {
        q.ring = (int*)malloc(sizeof(int) * (size_t)q.rcap);
        q.st = (uint8_t*)calloc((size_t)im.nmp, 1);
        q.holders = (int*)calloc((size_t)im.nmp, sizeof(int));
        if (!q.ring || !q.st || !q.holders) goto done;
    } else {
        q.rq = (int*)malloc(sizeof(int) * (size_t)q.cap);
        q.wq = (int*)malloc(sizeof(int) * (size_t)q.cap);
        q.last_round = (long long*)malloc(sizeof(long long) * (size_t)im.nmp);
        q.flag = (uint8_t*)calloc((size_t)im.nmp, 1);
        if (!q.rq || !q.wq || !q.last_round || !q.flag) goto done;
        for (int m = 0; m < im.nmp; ++m) q.last_round[m] = -1;
    }
    uint64_t rs = seed;
    for (;;) {
        const int r = phase(&im, &q, ws, workers, &rs, &S);
        if (r == 0) break;
        S.restarts++;  /* overflow: re-seed from the partial marker */
        if (S.restarts > 100000) goto done;
    }
    if (variant == QM_MUTEX) {
        for (int m = 0; m < im.nmp; ++m) S.state_left += q.st[m] != 0 || q.holders[m] != 0;
    } else if (variant == QM_UQ) {
        for (int m = 0; m < im.nmp; ++m) S.state_left += q.flag[m] != 0;
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint8_t* rj = (uint8_t*)marker + (size_t)y * pj;
            const uint16_t v = im.J[(size_t)y * w + x];
            if (dtype == 1) rj[x] = (uint8_t)v;
            else memcpy(rj + 2 * x, &v, 2);
        }
    rc = 0;
done:
    if (ws)
        for (int i = 0; i < workers; ++i) {
            free(ws[i].j0);
            free(ws[i].j);
            free(ws[i].i);
        }
    free(ws);
    free(im.J);
    free(im.I);
    free(q.ring);
    free(q.st);
    free(q.holders);
    free(q.rq);
    free(q.wq);
    free(q.last_round);
    free(q.flag);
    if (stats) *stats = S;
    return rc;
}

/* ------------------------------------------------------------- KATs ------ */
long qm_dq_fill_swap_kat(int g) {
    if (g <= 0) return -1;
    const long n = (long)g * g, cap = 8 * n;
    int* a = (int*)malloc(sizeof(int) * (size_t)cap);
    int* b = (int*)malloc(sizeof(int) * (size_t)cap);
    long* cnt = (long*)calloc((size_t)n, sizeof(long));
    if (!a || !b || !cnt) {
        free(a);
        free(b);
        free(cnt);
        return -2;
    }
    int *rq = a, *wq = b;
    long rsize = 0, wsize = 0;
    /* the reference's fill appends every coordinate to the read queue */
#define FILL()                                       \
    for (long i = 0; i < n; ++i) {                   \
        if (rsize < cap) rq[rsize] = (int)i;         \
        ++rsize;                                     \
    }
#define SWAP()                                       \
    do {                                             \
        if (wsize > cap) wsize = cap;                \
        int* t_ = rq;                                \
        rq = wq;                                     \
        wq = t_;                                     \
        rsize = wsize;                               \
        wsize = 0;                                   \
    } while (0)
    FILL() FILL() FILL() FILL()
    SWAP();
    FILL()
    SWAP();
    FILL() FILL()
#undef FILL
#undef SWAP
    for (long i = 0; i < MINV(rsize, cap); ++i) cnt[rq[i]]++;
    long bad = 0;
    for (long i = 0; i < n; ++i) bad += cnt[i] != 2;
    free(a);
    free(b);
    free(cnt);
    return bad;
}

long qm_uq_unique_kat(int g, int dup, uint64_t seed) {
    if (g <= 0 || dup <= 0) return -1;
    const long n = (long)g * g, len = n * dup;
    int* wq = (int*)malloc(sizeof(int) * (size_t)len);
    int* rq = (int*)malloc(sizeof(int) * (size_t)len);
    uint8_t* flags = (uint8_t*)calloc((size_t)n, 1);
    long* cnt = (long*)calloc((size_t)n, sizeof(long));
    if (!wq || !rq || !flags || !cnt) {
        free(wq);
        free(rq);
        free(flags);
        free(cnt);
        return -2;
    }
    for (long i = 0; i < len; ++i) wq[i] = (int)(i % n);
    uint64_t s = seed;
    for (long i = len - 1; i > 0; --i) {  /* Fisher-Yates */
        const long j = (long)(rng_next(&s) % (uint64_t)(i + 1));
        const int t = wq[i];
        wq[i] = wq[j];
        wq[j] = t;
    }
    long rsize = 0;
    for (long i = 0; i < len; ++i) {  /* extract uniques: test-and-set */
        const int m = wq[i];
        if (!flags[m]) {
            flags[m] = 1;
            rq[rsize++] = m;
        }
    }
    for (long i = 0; i < rsize; ++i) flags[rq[i]] = 0;  /* clear */
    long bad = 0;
    for (long i = 0; i < n; ++i) bad += flags[i] != 0;
    for (long i = 0; i < rsize; ++i) cnt[rq[i]]++;
    for (long i = 0; i < n; ++i) bad += cnt[i] != 1;
    free(wq);
    free(rq);
    free(flags);
    free(cnt);
    return bad;
}
