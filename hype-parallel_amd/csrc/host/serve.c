/*
 * serve.c -- request loop with continuous batching (host side, C).
 *
 * The reference solves one configured problem per program run
 * (hype.c:14-456).  The server accepts a stream of independent solve
 * requests, each with its own system, grid, end time and step budget, and
 * keeps as many of them resident in HBM as a byte budget allows.
 *
 * Rounds.  Each hype_server_poll prepares and enqueues one round: a number of
 * steps of the running instances (one persistent launch per RK stage per
 * kernel variant, replayed from a captured graph), followed by a copy of
 * every time state into pinned host memory.  Two rounds are kept in flight,
 * so the device never waits for the host: a poll waits only for the round
 * before the last one, and decides from that round's time states
 *   1. which instances have finished: their results are copied to the
 *      caller's buffers on a separate copy stream (a finished instance is
 *      no longer read or written by the steps still queued), and their slot
 *      and HBM block are freed;
 *   2. which queued requests to admit, while slots and the HBM budget allow,
 *      in arrival order or least work (cells x nvar x step budget) first;
 *      their initial states go up on the copy stream;
 *   3. which resident instances to advance: in arrival order, or least
 *      remaining work first, up to a cap on their cells (the others keep
 *      their state in HBM and wait).
 * Tile tables, descriptors and the new time states are staged in pinned
 * memory and copied on the compute stream, behind the round still running,
 * so they never change under a running step; the priming launches of the new
 * instances wait for their uploads.
 *
 * An instance's arithmetic does not depend on what else runs: results are
 * bitwise those of a solo hype_ctx solve.  Device states live in one arena
 * allocated at creation; blocks are placed first-fit and coalesced on
 * release, so admission never calls the device allocator.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "engine.h"

#define ARENA_ALIGN 256

typedef struct {
    size_t off, size;
} arena_blk;

typedef struct {
    char *base;
    size_t size, used;
    arena_blk *free_;       /* sorted by offset, coalesced */
    int nfree, capfree;
} arena;

typedef struct {
    hype_request r;
    int state;
    int slot;
    size_t off, bytes;
    double *inlet_h;        /* left-inlet table while resident (host copy), or NULL */
    double t_submit, t_start, t_done;
    hype_status st;
} srv_req;

struct hype_server {
    hype_server_opts o;
    hype_engine E;
    arena A;
    srv_req *req;
    int nreq, capreq;
    int *queue;             /* ids of queued requests, in arrival order */
    int nqueue;
    long long active_cells;
    int *slot_req;          /* request resident in each slot, or -1 */
    long long *slot_round;  /* first round that steps the slot's request */
    int *order;             /* scratch for the running-set selection */
    double *key;
    int *fresh;             /* slots admitted in this poll */
    int *clip;
    double *t0, *t_end;
    int *maxs;
    int active;
    double ms_per_step;     /* measured GPU time of one step of the mix */
    void *copy;             /* copy stream: initial states up, results down */
    void *ev_up;            /* uploads of this poll done */
    /* rounds: number enqueued so far, and per parity the pinned snapshot of
     * the time states after the round, its events and its step count */
    long long round;
    long long consumed;     /* rounds whose snapshot has been read */
    hype_dstate *snap[2];
    void *ev_snap[2], *ev0[2], *ev1[2];
    int steps[2];
    hype_dstate *d;         /* time states of the last consumed round */
    long long d_round;      /* that round's number (-1: none) */
    struct timespec t0c;
};

/* ---- arena ------------------------------------------------------------ */
static size_t round_up(size_t n) { return (n + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

static int arena_init(arena *A, size_t bytes)
{
    memset(A, 0, sizeof(*A));
    A->size = bytes / ARENA_ALIGN * ARENA_ALIGN;
    if (A->size == 0) return HYPE_E_INVALID;
    HYPE_TRY(hyped_malloc((void **)&A->base, A->size));
    A->capfree = 64;
    A->free_ = (arena_blk *)malloc((size_t)A->capfree * sizeof(arena_blk));
    if (!A->free_) return HYPE_E_NOMEM;
    A->free_[0].off = 0;
    A->free_[0].size = A->size;
    A->nfree = 1;
    return HYPE_OK;
}

static void arena_free_all(arena *A)
{
    hyped_free(A->base);
    free(A->free_);
    memset(A, 0, sizeof(*A));
}

/* first fit; returns 0 and the offset, or -1 if no block is large enough */
static int arena_alloc(arena *A, size_t bytes, size_t *off)
{
    int k;
    bytes = round_up(bytes);
    for (k = 0; k < A->nfree; ++k)
        if (A->free_[k].size >= bytes) {
            *off = A->free_[k].off;
            A->free_[k].off += bytes;
            A->free_[k].size -= bytes;
            if (A->free_[k].size == 0) {
                memmove(A->free_ + k, A->free_ + k + 1, (size_t)(A->nfree - k - 1) * sizeof(arena_blk));
                --A->nfree;
            }
            A->used += bytes;
            return 0;
        }
    return -1;
}

static int arena_release(arena *A, size_t off, size_t bytes)
{
    int k = 0;
    bytes = round_up(bytes);
    while (k < A->nfree && A->free_[k].off < off) ++k;
    if (A->nfree == A->capfree) {
        arena_blk *nb = (arena_blk *)realloc(A->free_, (size_t)A->capfree * 2 * sizeof(arena_blk));
        if (!nb) return HYPE_E_NOMEM;
        A->free_ = nb;
        A->capfree *= 2;
    }
    memmove(A->free_ + k + 1, A->free_ + k, (size_t)(A->nfree - k) * sizeof(arena_blk));
    A->free_[k].off = off;
    A->free_[k].size = bytes;
    ++A->nfree;
    A->used -= bytes;
    if (k + 1 < A->nfree && A->free_[k].off + A->free_[k].size == A->free_[k + 1].off) {
        A->free_[k].size += A->free_[k + 1].size;
        memmove(A->free_ + k + 1, A->free_ + k + 2, (size_t)(A->nfree - k - 2) * sizeof(arena_blk));
        --A->nfree;
    }
    if (k > 0 && A->free_[k - 1].off + A->free_[k - 1].size == A->free_[k].off) {
        A->free_[k - 1].size += A->free_[k].size;
        memmove(A->free_ + k, A->free_ + k + 1, (size_t)(A->nfree - k - 1) * sizeof(arena_blk));
        --A->nfree;
    }
    return HYPE_OK;
}

/* ---- server ----------------------------------------------------------- */
static double now_s(const hype_server *s)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - s->t0c.tv_sec) + 1e-9 * (double)(t.tv_nsec - s->t0c.tv_nsec);
}

static size_t state_bytes(const hype_problem *p)
{
    return (size_t)p->nx * p->ny * hype_nvar(p->system) * sizeof(double);
}

void hype_server_destroy(hype_server *s)
{
    int k;
    if (!s) return;
    if (s->copy) hyped_stream_sync(s->copy);
    for (k = 0; k < s->nreq; ++k) free(s->req[k].inlet_h);
    eng_free(&s->E);
    arena_free_all(&s->A);
    for (k = 0; k < 2; ++k) {
        hyped_host_free(s->snap[k]);
        if (s->ev_snap[k]) hyped_event_destroy(s->ev_snap[k]);
        if (s->ev0[k]) hyped_event_destroy(s->ev0[k]);
        if (s->ev1[k]) hyped_event_destroy(s->ev1[k]);
    }
    if (s->ev_up) hyped_event_destroy(s->ev_up);
    if (s->copy) hyped_stream_destroy(s->copy);
    free(s->req);
    free(s->queue);
    free(s->slot_req);
    free(s->slot_round);
    free(s->order);
    free(s->key);
    free(s->fresh);
    free(s->clip);
    free(s->t0);
    free(s->t_end);
    free(s->maxs);
    free(s->d);
    free(s);
}

int hype_server_create(const hype_server_opts *opts, hype_server **out)
{
    hype_server *s;
    size_t budget;
    int k, rc, cap;
    if (!out) return HYPE_E_INVALID;
    *out = NULL;
    s = (hype_server *)calloc(1, sizeof(*s));
    if (!s) return HYPE_E_NOMEM;
    if (opts) s->o = *opts;
    if (s->o.max_active < 1) s->o.max_active = 512;
    if (s->o.round_ms <= 0.0) s->o.round_ms = 20.0;
    if (s->o.policy != HYPE_SERVE_FIFO && s->o.policy != HYPE_SERVE_SHORTEST) {
        free(s);
        return HYPE_E_INVALID;
    }
    clock_gettime(CLOCK_MONOTONIC, &s->t0c);
    budget = s->o.hbm_budget_bytes;
    if ((rc = hyped_init()) != HYPE_OK) {
        free(s);
        return rc;
    }
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if ((rc = hyped_mem_info(&fr, &tot)) != HYPE_OK) {
            free(s);
            return rc;
        }
        budget = fr / 4 * 3;
    }
    s->o.hbm_budget_bytes = budget;
    if (s->o.max_active_cells == 0) {
        /* tiles of a few hundred rows for every resident wave of the device:
         * enough for short halo overhead, more only delays everyone */
        int ncu = hyped_num_cus();
        s->o.max_active_cells = (long long)(ncu > 0 ? ncu : 256) * 4 * HYPE_TILE_W * 256;
    }
    cap = s->o.max_active;
    s->slot_req = (int *)malloc((size_t)cap * sizeof(int));
    s->slot_round = (long long *)calloc((size_t)cap, sizeof(long long));
    s->order = (int *)malloc((size_t)cap * sizeof(int));
    s->key = (double *)malloc((size_t)cap * sizeof(double));
    s->fresh = (int *)malloc((size_t)cap * sizeof(int));
    s->clip = (int *)malloc((size_t)cap * sizeof(int));
    s->t0 = (double *)malloc((size_t)cap * sizeof(double));
    s->t_end = (double *)malloc((size_t)cap * sizeof(double));
    s->maxs = (int *)malloc((size_t)cap * sizeof(int));
    s->d = (hype_dstate *)calloc((size_t)cap, sizeof(hype_dstate));
    if (!s->slot_req || !s->slot_round || !s->order || !s->key || !s->fresh || !s->clip || !s->t0 ||
        !s->t_end || !s->maxs || !s->d) {
        hype_server_destroy(s);
        return HYPE_E_NOMEM;
    }
    for (k = 0; k < cap; ++k) s->slot_req[k] = -1;
    s->d_round = -1;
    if ((rc = eng_init(&s->E, cap, s->o.tile_rows, ENG_ASYNC)) != HYPE_OK ||
        (rc = arena_init(&s->A, budget)) != HYPE_OK || (rc = hyped_stream_create(&s->copy)) != HYPE_OK ||
        (rc = hyped_event_create_sync(&s->ev_up)) != HYPE_OK) {
        hype_server_destroy(s);
        return rc;
    }
    for (k = 0; k < 2 && rc == HYPE_OK; ++k) {
        if ((rc = hyped_host_alloc((void **)&s->snap[k], (size_t)cap * sizeof(hype_dstate))) != HYPE_OK ||
            (rc = hyped_event_create_sync(&s->ev_snap[k])) != HYPE_OK ||
            (rc = hyped_event_create(&s->ev0[k])) != HYPE_OK || (rc = hyped_event_create(&s->ev1[k])) != HYPE_OK)
            break;
    }
    if (rc != HYPE_OK) {
        hype_server_destroy(s);
        return rc;
    }
    *out = s;
    return HYPE_OK;
}

double hype_server_now(hype_server *s) { return s ? now_s(s) : 0.0; }

size_t hype_server_bytes_in_use(hype_server *s) { return s ? s->A.used : 0; }

int hype_server_submit(hype_server *s, const hype_request *r, int *id)
{
    int nv;
    srv_req *q;
    if (!s || !r || !r->U0) return HYPE_E_INVALID;
    HYPE_TRY(hype_check_problem(&r->problem, &nv));
    if (r->problem.order < 1) return HYPE_E_UNSUPPORTED;
    if (round_up(3 * state_bytes(&r->problem) + hype_inlet_doubles(&r->problem) * sizeof(double)) > s->A.size)
        return HYPE_E_NOMEM; /* never admissible */
    if (s->nreq == s->capreq) {
        int cap = s->capreq ? 2 * s->capreq : 256;
        srv_req *nr = (srv_req *)realloc(s->req, (size_t)cap * sizeof(srv_req));
        int *nq;
        if (!nr) return HYPE_E_NOMEM;
        s->req = nr;
        nq = (int *)realloc(s->queue, (size_t)cap * sizeof(int));
        if (!nq) return HYPE_E_NOMEM;
        s->queue = nq;
        s->capreq = cap;
    }
    q = &s->req[s->nreq];
    memset(q, 0, sizeof(*q));
    q->r = *r;
    q->state = HYPE_REQ_QUEUED;
    q->slot = -1;
    q->bytes = 3 * state_bytes(&r->problem) + hype_inlet_doubles(&r->problem) * sizeof(double);
    q->inlet_h = NULL;
    q->t_submit = now_s(s);
    if (id) *id = s->nreq;
    s->queue[s->nqueue++] = s->nreq;
    ++s->nreq;
    return HYPE_OK;
}

/* queue position of the next request to admit */
static int pick(const hype_server *s)
{
    int k, best = 0;
    double bw = 0.0;
    if (s->o.policy == HYPE_SERVE_FIFO) return 0;
    for (k = 0; k < s->nqueue; ++k) {
        const srv_req *q = &s->req[s->queue[k]];
        double w = (double)q->r.problem.nx * q->r.problem.ny * hype_nvar(q->r.problem.system) * q->r.max_steps;
        if (k == 0 || w < bw) {
            bw = w;
            best = k;
        }
    }
    return best;
}

/* the last consumed time state describes the slot's current request */
static int state_valid(const hype_server *s, int slot)
{
    return s->d_round >= 0 && s->d_round >= s->slot_round[slot];
}

static double remaining_work(const hype_server *s, int slot)
{
    const srv_req *q = &s->req[s->slot_req[slot]];
    const hype_dstate *d = &s->d[slot];
    double steps = q->r.max_steps;
    if (state_valid(s, slot)) {
        steps = (double)(d->max_steps - d->step);
        if (d->dt > 0.0 && d->t_end < 1e300) {
            const double by_time = (d->t_end - d->t) / d->dt + 1.0;
            if (by_time < steps) steps = by_time;
        }
    }
    return (double)q->r.problem.nx * q->r.problem.ny * hype_nvar(q->r.problem.system) * steps;
}

static int cmp_key(const void *a, const void *b, void *ctx)
{
    const double *key = (const double *)ctx;
    const int x = *(const int *)a, y = *(const int *)b;
    if (key[x] < key[y]) return -1;
    if (key[x] > key[y]) return 1;
    return x - y;
}

/* choose the running subset; returns 1 if it changed */
static int select_running(hype_server *s)
{
    int k, n = 0, changed = 0;
    long long cells = 0;
    for (k = 0; k < s->o.max_active; ++k) {
        if (s->slot_req[k] < 0) continue;
        s->order[n++] = k;
        s->key[k] = s->o.policy == HYPE_SERVE_SHORTEST ? remaining_work(s, k) : (double)s->slot_req[k];
    }
    qsort_r(s->order, (size_t)n, sizeof(int), cmp_key, s->key);
    for (k = 0; k < n; ++k) {
        const int slot = s->order[k];
        const srv_req *q = &s->req[s->slot_req[slot]];
        const long long c = (long long)q->r.problem.nx * q->r.problem.ny;
        const int on = k == 0 || s->o.max_active_cells < 0 || cells + c <= s->o.max_active_cells;
        if (on) cells += c;
        if (on != s->E.run[slot]) {
            eng_set_running(&s->E, slot, on);
            changed = 1;
        }
    }
    return changed;
}

static int finished(const hype_dstate *d, unsigned long long *err)
{
    unsigned long long e = d->step > 0 ? d->err[(d->step - 1) & 1] : d->err[1];
    *err = e;
    return d->done || d->step >= d->max_steps || !(d->t < d->t_end) || e != ~0ull;
}

/* Retire the finished request of a slot.  Its state is no longer touched by
 * the steps still queued (they return at once for a finished instance), so
 * the result is copied on the copy stream while they run; the slot and the
 * HBM block are free once the copy has landed: later writes to the slot's
 * descriptor and time state are ordered behind the queued steps, and later
 * uploads to the block behind this copy. */
static int retire(hype_server *s, int slot)
{
    srv_req *q = &s->req[s->slot_req[slot]];
    const hype_dstate *d = &s->d[slot];
    unsigned long long e;
    finished(d, &e);
    if (q->r.U_out) {
        HYPE_TRY(hyped_memcpy_d2h(q->r.U_out, s->A.base + q->off, state_bytes(&q->r.problem), s->copy));
        HYPE_TRY(hyped_stream_sync(s->copy));
    }
    q->st.t = d->t;
    q->st.dt = d->dt;
    q->st.steps = d->step;
    q->st.done = 1;
    q->st.err = e != ~0ull ? HYPE_E_INADMISSIBLE : HYPE_OK;
    q->t_done = now_s(s);
    q->state = HYPE_REQ_DONE;
    q->slot = -1;
    HYPE_TRY(eng_detach(&s->E, slot));
    HYPE_TRY(arena_release(&s->A, q->off, q->bytes));
    free(q->inlet_h);
    q->inlet_h = NULL;
    s->slot_req[slot] = -1;
    s->active_cells -= (long long)q->r.problem.nx * q->r.problem.ny;
    --s->active;
    return HYPE_OK;
}

/* read the time states after round k (waits for it) */
static int consume(hype_server *s, long long k)
{
    const int p = (int)(k & 1);
    float ms = 0.0f;
    HYPE_TRY(hyped_event_sync(s->ev_snap[p]));
    memcpy(s->d, s->snap[p], (size_t)s->o.max_active * sizeof(hype_dstate));
    s->d_round = k;
    s->consumed = k + 1;
    if (s->steps[p] > 0 && hyped_event_elapsed_ms(s->ev0[p], s->ev1[p], &ms) == HYPE_OK && ms > 0.0f)
        s->ms_per_step = (double)ms / s->steps[p];
    return HYPE_OK;
}

int hype_server_poll(hype_server *s, int *active, int *queued)
{
    int slot, changed = 0, nfresh = 0, rc = HYPE_OK, k;
    double t;
    if (!s) return HYPE_E_INVALID;
    /* the round before the last one; the device keeps the last.  With no
     * instance left, the last one too, so that nothing stays unread. */
    if (s->round - s->consumed >= 2 || (s->active == 0 && s->round > s->consumed))
        if ((rc = consume(s, s->active == 0 ? s->round - 1 : s->round - 2)) != HYPE_OK) goto out;
    /* 1: retire what had finished by then */
    if (s->active > 0 && s->d_round >= 0) {
        for (slot = 0; slot < s->o.max_active; ++slot) {
            unsigned long long e;
            if (s->slot_req[slot] < 0 || !state_valid(s, slot) || !finished(&s->d[slot], &e)) continue;
            if ((rc = retire(s, slot)) != HYPE_OK) goto out;
            changed = 1;
        }
    }
    /* 2: admit while a slot and HBM are free; initial states go up on the
     * copy stream while the queued steps run */
    t = now_s(s);
    while (s->nqueue > 0) {
        int kq = pick(s), id = s->queue[kq];
        srv_req *q = &s->req[id];
        long long cells = (long long)q->r.problem.nx * q->r.problem.ny;
        double *U;
        size_t m;
        for (slot = 0; slot < s->o.max_active && s->slot_req[slot] >= 0; ++slot) {
        }
        if (slot == s->o.max_active) break;
        if (arena_alloc(&s->A, q->bytes, &q->off) != 0) break;
        U = (double *)(s->A.base + q->off);
        m = (size_t)cells * hype_nvar(q->r.problem.system);
        {
            const size_t ni = hype_inlet_doubles(&q->r.problem);
            double *inlet = NULL;
            if (ni > 0) {  /* the table's host copy lives until the request retires */
                q->inlet_h = (double *)malloc(ni * sizeof(double));
                if (!q->inlet_h) {
                    arena_release(&s->A, q->off, q->bytes);
                    rc = HYPE_E_NOMEM;
                    goto out;
                }
                hype_inlet_table(&q->r.problem, q->inlet_h);
                inlet = U + 3 * m;
                rc = hyped_memcpy_h2d(inlet, q->inlet_h, ni * sizeof(double), s->copy);
            }
            if (rc != HYPE_OK || (rc = hyped_memcpy_h2d(U, q->r.U0, m * sizeof(double), s->copy)) != HYPE_OK ||
                (rc = eng_attach(&s->E, slot, &q->r.problem, U, U + m, U + 2 * m, inlet)) != HYPE_OK) {
                arena_release(&s->A, q->off, q->bytes);
                free(q->inlet_h);
                q->inlet_h = NULL;
                goto out;
            }
        }
        q->slot = slot;
        q->state = HYPE_REQ_RUNNING;
        q->t_start = t;
        s->slot_req[slot] = id;
        s->slot_round[slot] = s->round; /* primed and first stepped in the round enqueued below */
        s->fresh[nfresh] = slot;
        s->clip[nfresh] = (q->r.flags & HYPE_TS_CLIP) != 0;
        s->t0[nfresh] = q->r.t0;
        s->t_end[nfresh] = q->r.t_end;
        s->maxs[nfresh] = q->r.max_steps;
        ++nfresh;
        ++s->active;
        s->active_cells += cells;
        memmove(s->queue + kq, s->queue + kq + 1, (size_t)(s->nqueue - kq - 1) * sizeof(int));
        --s->nqueue;
        changed = 1;
    }
    if (s->active == 0) goto out;
    /* 3: the instances advanced this round */
    if (select_running(s)) changed = 1;
    /* 4: tables, descriptors and new time states, staged for this round and
     * copied behind the queued steps; the priming waits for the uploads */
    eng_flip(&s->E);
    if (changed) {
        if (s->o.tile_rows < 1) eng_set_rows(&s->E, eng_pick_rows(&s->E));
        if ((rc = eng_commit(&s->E)) != HYPE_OK) goto out;
        if (nfresh > 0) {
            if ((rc = hyped_event_record(s->ev_up, s->copy)) != HYPE_OK) goto out;
            if ((rc = hyped_stream_wait(s->E.stream, s->ev_up)) != HYPE_OK) goto out;
            if ((rc = eng_start(&s->E, s->fresh, nfresh, s->t0, s->t_end, s->maxs, s->clip, NULL, NULL)) != HYPE_OK)
                goto out;
        }
    }
    /* 5: the round's steps, then the snapshot of the time states */
    {
        const int p = (int)(s->round & 1);
        int n = s->o.steps_per_round;
        if (n < 1) {
            n = s->ms_per_step > 0.0 ? (int)(s->o.round_ms / s->ms_per_step) : 4;
            if (n < 1) n = 1;
            if (n > 256) n = 256;
        }
        if ((rc = hyped_event_record(s->ev0[p], s->E.stream)) != HYPE_OK) goto out;
        if ((rc = eng_step(&s->E, n)) != HYPE_OK) goto out;
        if ((rc = hyped_event_record(s->ev1[p], s->E.stream)) != HYPE_OK) goto out;
        if ((rc = hyped_memcpy_d2h(s->snap[p], s->E.st_d, (size_t)s->o.max_active * sizeof(hype_dstate),
                                   s->E.stream)) != HYPE_OK)
            goto out;
        if ((rc = hyped_event_record(s->ev_snap[p], s->E.stream)) != HYPE_OK) goto out;
        s->steps[p] = n;
        ++s->round;
    }
out:
    for (k = 0; k < nfresh; ++k) s->fresh[k] = -1;
    if (active) *active = s->active;
    if (queued) *queued = s->nqueue;
    return rc;
}

int hype_server_drain(hype_server *s)
{
    int a = 0, q = 0;
    if (!s) return HYPE_E_INVALID;
    do {
        HYPE_TRY(hype_server_poll(s, &a, &q));
    } while (a > 0 || q > 0);
    return HYPE_OK;
}

int hype_server_info(hype_server *s, int id, hype_request_info *info)
{
    const srv_req *q;
    if (!s || !info) return HYPE_E_INVALID;
    if (id < 0 || id >= s->nreq) return HYPE_E_NOTFOUND;
    q = &s->req[id];
    memset(info, 0, sizeof(*info));
    info->state = q->state;
    info->status = q->st;
    info->t_submit = q->t_submit;
    info->t_start = q->t_start;
    info->t_done = q->t_done;
    info->bytes = q->bytes;
    return HYPE_OK;
}
