/* engine.c -- multi-instance stepping engine (see engine.h). */
#include "engine.h"

#include <stdlib.h>
#include <string.h>

#define GROUP_TILE_CAP0 1024

int eng_init(hype_engine *E, int cap, int target_rows, unsigned flags)
{
    int p;
    memset(E, 0, sizeof(*E));
    if (cap < 1) return HYPE_E_INVALID;
    HYPE_TRY(hyped_init());
    E->cap = cap;
    E->layout = -1;
    E->target_rows = target_rows;
    E->flags = flags;
    E->ncu = hyped_num_cus();
    if (E->ncu < 1) E->ncu = 256;
    E->inst_h = (hype_inst *)calloc((size_t)cap, sizeof(hype_inst));
    E->P = (hype_problem *)calloc((size_t)cap, sizeof(hype_problem));
    E->used = (int *)calloc((size_t)cap, sizeof(int));
    E->group_of = (int *)calloc((size_t)cap, sizeof(int));
    E->nsy = (int *)calloc((size_t)cap, sizeof(int));
    E->fresh = (int *)calloc((size_t)cap, sizeof(int));
    E->run = (int *)calloc((size_t)cap, sizeof(int));
    if (!E->inst_h || !E->P || !E->used || !E->group_of || !E->nsy || !E->fresh || !E->run) return HYPE_E_NOMEM;
    HYPE_TRY(hyped_malloc((void **)&E->inst_d, (size_t)cap * sizeof(hype_inst)));
    HYPE_TRY(hyped_malloc((void **)&E->st_d, (size_t)cap * sizeof(hype_dstate)));
    for (p = 0; p < 2; ++p) {
        HYPE_TRY(hyped_host_alloc((void **)&E->inst_stage[p], (size_t)cap * sizeof(hype_inst)));
        HYPE_TRY(hyped_host_alloc((void **)&E->st_stage[p], (size_t)cap * sizeof(hype_dstate)));
    }
    HYPE_TRY(hyped_stream_create(&E->stream));
    HYPE_TRY(hyped_event_create_sync(&E->ev_fork));
    /* time states start zeroed: max_steps = 0 and no history pointer, so a
     * stage launch over a slot that was never started returns at once */
    HYPE_TRY(hyped_fill_u64(E->st_d, 0ull, (size_t)cap * sizeof(hype_dstate) / 8, E->stream));
    return hyped_stream_sync(E->stream);
}

void eng_free(hype_engine *E)
{
    int g;
    if (E->stream) hyped_stream_sync(E->stream);
    if (E->graph) hyped_graph_destroy(E->graph);
    for (g = 0; g < E->ng; ++g) {
        eng_group *G = &E->g[g];
        int p;
        for (p = 0; p < 2; ++p) {
            hyped_host_free(G->stage_tiles[p]);
            hyped_host_free(G->stage_prime[p]);
        }
        hyped_free(G->tiles_d);
        hyped_free(G->queue_d);
        hyped_free(G->ntiles_d);
        hyped_free(G->prime_d);
        hyped_free(G->nprime_d);
        if (G->ev_done) hyped_event_destroy(G->ev_done);
        if (G->stream) hyped_stream_destroy(G->stream);
    }
    hyped_free(E->inst_d);
    hyped_free(E->st_d);
    for (g = 0; g < 2; ++g) {
        hyped_host_free(E->inst_stage[g]);
        hyped_host_free(E->st_stage[g]);
    }
    if (E->ev_fork) hyped_event_destroy(E->ev_fork);
    if (E->stream) hyped_stream_destroy(E->stream);
    free(E->inst_h);
    free(E->P);
    free(E->used);
    free(E->group_of);
    free(E->nsy);
    free(E->fresh);
    free(E->run);
    memset(E, 0, sizeof(*E));
}

static int same_variant(const hype_variant *a, const hype_variant *b)
{
    return a->system == b->system && a->prim == b->prim && a->order == b->order && a->fma == b->fma &&
           a->pair == b->pair && a->inlet == b->inlet;
}

static int find_group(hype_engine *E, const hype_variant *v)
{
    int g;
    for (g = 0; g < E->ng; ++g)
        if (same_variant(&E->g[g].v, v)) return g;
    if (E->ng >= HYPE_MAX_GROUPS) return -1;
    {
        eng_group *G = &E->g[E->ng];
        memset(G, 0, sizeof(*G));
        G->v = *v;
        G->nstages = hyped_num_stages(v->order);
        G->tile_cap = GROUP_TILE_CAP0;
        G->prime_cap = GROUP_TILE_CAP0;
        if (hyped_malloc((void **)&G->tiles_d, (size_t)G->tile_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_malloc((void **)&G->prime_d, (size_t)G->tile_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_malloc((void **)&G->ntiles_d, 2 * sizeof(int)) != HYPE_OK ||
            hyped_malloc((void **)&G->nprime_d, 2 * sizeof(int)) != HYPE_OK ||
            hyped_fill_u64(G->ntiles_d, 0ull, 1, E->stream) != HYPE_OK ||
            hyped_fill_u64(G->nprime_d, 0ull, 1, E->stream) != HYPE_OK ||
            hyped_stream_create(&G->stream) != HYPE_OK || hyped_event_create_sync(&G->ev_done) != HYPE_OK ||
            hyped_host_alloc((void **)&G->stage_tiles[0], (size_t)G->tile_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_host_alloc((void **)&G->stage_tiles[1], (size_t)G->tile_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_host_alloc((void **)&G->stage_prime[0], (size_t)G->prime_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_host_alloc((void **)&G->stage_prime[1], (size_t)G->prime_cap * sizeof(hype_tile)) != HYPE_OK ||
            hyped_malloc((void **)&G->queue_d, 4 * sizeof(unsigned)) != HYPE_OK ||
            hyped_fill_u64(G->queue_d, 0ull, 2, E->stream) != HYPE_OK)
            return -1;
        G->grid = E->ncu * hyped_blocks_per_cu(v);
        if (G->grid < 1) G->grid = E->ncu;
        G->bwaves = hyped_block_waves(v);
        /* a new branch: the step graph must be captured again (once the
         * replays already queued have run) */
        if (E->graph) {
            if (hyped_stream_sync(E->stream) != HYPE_OK) return -1;
            hyped_graph_destroy(E->graph);
            E->graph = NULL;
        }
        return E->ng++;
    }
}

/* the kernel variant of a problem in this engine */
static int slot_variant(const hype_engine *E, const hype_problem *p, hype_variant *v)
{
    const int modes = hype_layouts(p, hyped_pair_modes(p->system));
    if (!modes) return HYPE_E_UNSUPPORTED;
    v->system = p->system;
    v->prim = p->recons_prim;
    v->order = p->order;
    v->fma = (p->options & HYPE_OPT_FMA) != 0;
    /* split tiles where the system has no other layout (or the problem asks
     * for them), or where both are built and the engine prefers them; by
     * default a batch keeps many waves per SIMD busy, where one-wave tiles
     * hide latency better */
    if (modes == 3 && E->layout >= 0) v->pair = E->layout;
    else v->pair = (modes & 1) ? 0 : 1;
    v->inlet = p->bc[0] == HYPE_BC_INLET;
    return HYPE_OK;
}

int eng_attach(hype_engine *E, int slot, const hype_problem *p, double *U, double *A, double *B,
               const double *inlet)
{
    hype_variant v;
    int g, nv = 0;
    if (slot < 0 || slot >= E->cap || E->used[slot]) return HYPE_E_INVALID;
    HYPE_TRY(hype_check_problem(p, &nv));
    if (p->order < 1) return HYPE_E_UNSUPPORTED; /* forward Euler stepping stays single-instance */
    HYPE_TRY(slot_variant(E, p, &v));
    g = find_group(E, &v);
    if (g < 0) return HYPE_E_DEVICE;
    E->P[slot] = *p;
    E->group_of[slot] = g;
    E->used[slot] = 1;
    E->run[slot] = 1;
    E->fresh[slot] = 1;
    hype_fill_inst(&E->inst_h[slot], p, 1);
    E->inst_h[slot].buf[0] = U;
    E->inst_h[slot].buf[1] = A;
    E->inst_h[slot].buf[2] = B;
    E->inst_h[slot].buf[3] = NULL;
    E->inst_h[slot].st = E->st_d + slot;
    E->inst_h[slot].inlet = inlet;
    return HYPE_OK;
}

int eng_detach(hype_engine *E, int slot)
{
    if (slot < 0 || slot >= E->cap || !E->used[slot]) return HYPE_E_INVALID;
    E->used[slot] = 0;
    E->run[slot] = 0;
    E->fresh[slot] = 0;
    return HYPE_OK;
}

int eng_set_layout(hype_engine *E, int layout)
{
    int s;
    if (layout < -1 || layout > 1) return HYPE_E_INVALID;
    E->layout = layout;
    for (s = 0; s < E->cap; ++s) {
        hype_variant v;
        int g;
        if (!E->used[s]) continue;
        HYPE_TRY(slot_variant(E, &E->P[s], &v));
        g = find_group(E, &v);
        if (g < 0) return HYPE_E_DEVICE;
        E->group_of[s] = g;
    }
    return HYPE_OK;
}

int eng_set_running(hype_engine *E, int slot, int on)
{
    if (slot < 0 || slot >= E->cap || !E->used[slot]) return HYPE_E_INVALID;
    E->run[slot] = on ? 1 : 0;
    return HYPE_OK;
}

static int rows_for(const hype_engine *E, int slot)
{
    int ny = E->P[slot].ny, h = E->target_rows > 0 ? E->target_rows : 16;
    int nsy = (ny + h - 1) / h;
    if (nsy < 1) nsy = 1;
    if (nsy > ny) nsy = ny;
    return nsy;
}

/* tiles of the used slots of group g (g < 0: all groups) */
static int count_tiles(const hype_engine *E, int g, int only_fresh)
{
    int s, t = 0;
    for (s = 0; s < E->cap; ++s)
        if (E->used[s] && (g < 0 || E->group_of[s] == g) && (only_fresh ? E->fresh[s] : E->run[s]))
            t += hype_strips(E->P[s].nx) * rows_for(E, s);
    return t;
}

int eng_total_tiles(hype_engine *E) { return count_tiles(E, -1, 0); }

/* grow the stepping table of a group to hold n tiles; it is a kernel
 * argument of the captured graph, so a new one invalidates the graph.  The
 * caller rewrites the whole table afterwards (eng_commit). */
static int ensure_step_cap(hype_engine *E, eng_group *G, int n)
{
    int cap = G->tile_cap, p;
    if (n <= cap) return HYPE_OK;
    while (cap < n) cap *= 2;
    HYPE_TRY(hyped_stream_sync(E->stream));
    hyped_free(G->tiles_d);
    G->tiles_d = NULL;
    G->tile_cap = 0;
    G->ntiles = 0;
    HYPE_TRY(hyped_fill_u64(G->ntiles_d, 0ull, 1, E->stream));
    HYPE_TRY(hyped_malloc((void **)&G->tiles_d, (size_t)cap * sizeof(hype_tile)));
    for (p = 0; p < 2; ++p) {
        hyped_host_free(G->stage_tiles[p]);
        G->stage_tiles[p] = NULL;
        HYPE_TRY(hyped_host_alloc((void **)&G->stage_tiles[p], (size_t)cap * sizeof(hype_tile)));
    }
    G->tile_cap = cap;
    if (E->graph) {
        hyped_graph_destroy(E->graph);
        E->graph = NULL;
    }
    return HYPE_OK;
}

/* grow the priming table of a group; the priming launch is not captured,
 * and the stepping table and its length are left as they are */
static int ensure_prime_cap(hype_engine *E, eng_group *G, int n)
{
    int cap = G->prime_cap, p;
    if (n <= cap) return HYPE_OK;
    while (cap < n) cap *= 2;
    HYPE_TRY(hyped_stream_sync(E->stream));
    hyped_free(G->prime_d);
    G->prime_d = NULL;
    G->prime_cap = 0;
    HYPE_TRY(hyped_malloc((void **)&G->prime_d, (size_t)cap * sizeof(hype_tile)));
    for (p = 0; p < 2; ++p) {
        hyped_host_free(G->stage_prime[p]);
        G->stage_prime[p] = NULL;
        HYPE_TRY(hyped_host_alloc((void **)&G->stage_prime[p], (size_t)cap * sizeof(hype_tile)));
    }
    G->prime_cap = cap;
    return HYPE_OK;
}

/* Common tile height for the running slots: the tile-shape model of
 * hype_pick_segments applied to the union of all tiles, with the smallest
 * occupancy among the groups. */
int eng_pick_rows(hype_engine *E)
{
    int h, best = 16, g, simds = 4 * E->ncu, occ = 1 << 30;
    double best_cost = 1e300;
    for (g = 0; g < E->ng; ++g) {
        int o = E->g[g].grid * E->g[g].bwaves / simds;
        if (o < occ) occ = o;
    }
    if (occ < 1 || occ == 1 << 30) occ = 1;
    for (h = 1; h <= 128; ++h) {
        long waves = 0;
        int s;
        double per = 0.0;
        for (s = 0; s < E->cap; ++s)
            if (E->used[s] && E->run[s]) {
                const int ny = E->P[s].ny, nsy = (ny + h - 1) / h, ht = (ny + nsy - 1) / nsy;
                const int bw = E->g[E->group_of[s]].bwaves;
                const double pw = hype_tile_rows(ht, bw > 1);
                waves += (long)hype_strips(E->P[s].nx) * nsy * bw;
                if (pw > per) per = pw;
            }
        if (waves == 0) break;
        {
            const double cost = hype_tile_cost(waves, per, simds, occ);
            if (cost < best_cost - 1e-9) {
                best_cost = cost;
                best = h;
            }
        }
    }
    return best;
}

int eng_set_rows(hype_engine *E, int rows)
{
    if (rows < 1) return HYPE_E_INVALID;
    E->target_rows = rows;
    return HYPE_OK;
}

static int cmp_tile_h(const void *a, const void *b)
{
    const hype_tile *x = (const hype_tile *)a, *y = (const hype_tile *)b;
    if (x->h != y->h) return y->h - x->h; /* taller first: longest-job-first queue */
    if (x->inst != y->inst) return x->inst - y->inst;
    if (x->y0 != y->y0) return x->y0 - y->y0;
    return x->x0 - y->x0;
}

/* Tiles of the group's slots (all running ones, or the fresh ones), tallest
 * first: the persistent launch's workgroups pull them in this order, so the
 * launch ends on the shortest tiles.  A tile takes a sizeable part of a
 * launch (the batched set runs about eight tiles per workgroup, with heights
 * that differ by a third), and this order was measured 2.5 % faster than
 * any order that gives each XCD a band of neighbouring tiles for L2 reuse. */
static int build_tiles(hype_engine *E, int g, int only_fresh, hype_tile *buf, int cap)
{
    int s, n = 0;
    for (s = 0; s < E->cap; ++s) {
        int k, m;
        if (!E->used[s] || E->group_of[s] != g) continue;
        if (only_fresh ? !E->fresh[s] : !E->run[s]) continue;
        E->nsy[s] = rows_for(E, s);
        E->inst_h[s].nsy = E->nsy[s];
        m = hype_make_tiles(s, E->P[s].nx, E->P[s].ny, hype_strips(E->P[s].nx), E->nsy[s], buf + n, cap - n);
        if (m < 0) return -1;
        for (k = 0; k < m; ++k) buf[n + k].flags = (k == 0) ? 1 : 0;
        n += m;
    }
    qsort(buf, (size_t)n, sizeof(hype_tile), cmp_tile_h);
    return n;
}

void eng_flip(hype_engine *E) { E->par ^= 1; }

int eng_commit(hype_engine *E)
{
    int g, rc = HYPE_OK;
    for (g = 0; g < E->ng && rc == HYPE_OK; ++g) {
        eng_group *G = &E->g[g];
        hype_tile *buf;
        int n;
        if ((rc = ensure_step_cap(E, G, count_tiles(E, g, 0))) != HYPE_OK) break;
        buf = G->stage_tiles[E->par];
        n = build_tiles(E, g, 0, buf, G->tile_cap);
        if (n < 0) {
            rc = HYPE_E_INVALID; /* a grid the tiler cannot split */
            break;
        }
        G->ntiles = n;
        if (n > 0) rc = hyped_memcpy_h2d(G->tiles_d, buf, (size_t)n * sizeof(hype_tile), E->stream);
        if (rc == HYPE_OK) rc = hyped_fill_u64(G->ntiles_d, (unsigned long long)(unsigned)n, 1, E->stream);
    }
    if (rc == HYPE_OK) {
        memcpy(E->inst_stage[E->par], E->inst_h, (size_t)E->cap * sizeof(hype_inst));
        rc = hyped_memcpy_h2d(E->inst_d, E->inst_stage[E->par], (size_t)E->cap * sizeof(hype_inst), E->stream);
    }
    if (rc == HYPE_OK && !(E->flags & ENG_ASYNC)) rc = hyped_stream_sync(E->stream);
    return rc;
}

/* The persistent launch of a group over its stepping or priming table: the
 * table's length is read from device memory, so the tables may change
 * between graph replays; the last workgroup to finish resets the queue. */
static hype_launch base(hype_engine *E, eng_group *G, int mode, int stage)
{
    hype_launch L;
    memset(&L, 0, sizeof(L));
    L.v = G->v;
    L.insts = E->inst_d;
    L.mode = mode;
    L.stage = stage;
    L.grid = G->grid;
    if (mode == HYPE_MODE_PRIME) {
        L.tiles = G->prime_d;
        L.ntiles_dev = G->nprime_d;
        L.queue = G->queue_d + 2;
    } else {
        L.tiles = G->tiles_d;
        L.ntiles_dev = G->ntiles_d;
        L.queue = G->queue_d;
    }
    L.ntiles = 1; /* launch even when the device-resident count is 0 */
    return L;
}

int eng_start(hype_engine *E, const int *slots, int n, const double *t0, const double *t_end,
              const int *max_steps, const int *clip, double **hist, const int *hist_len)
{
    hype_dstate *d = E->st_stage[E->par];
    int k, g, m, rc = HYPE_OK;
    for (k = 0; k < n; ++k)
        if (slots[k] < 0 || slots[k] >= E->cap || !E->used[slots[k]]) return HYPE_E_INVALID;
    if (n > E->cap) return HYPE_E_INVALID;
    /* the new time states, staged in pinned memory */
    memset(d, 0, (size_t)n * sizeof(hype_dstate));
    for (k = 0; k < n && rc == HYPE_OK; ++k) {
        const int s = slots[k];
        d[k].t = t0 ? t0[k] : 0.0;
        d[k].t_end = t_end[k];
        d[k].err[0] = d[k].err[1] = d[k].err_rhs = ~0ull;
        d[k].max_steps = max_steps[k];
        d[k].clip = clip ? (clip[k] != 0) : 0;
        d[k].dt_hist = hist ? hist[k] : NULL;
        d[k].hist_len = (hist && hist_len) ? hist_len[k] : 0;
        rc = hyped_memcpy_h2d(E->st_d + s, &d[k], sizeof(hype_dstate), E->stream);
        E->fresh[s] = 2; /* marked for priming */
    }
    if (rc != HYPE_OK) return rc;
    for (k = 0; k < E->cap; ++k) E->fresh[k] = (E->fresh[k] == 2);
    for (g = 0; g < E->ng && rc == HYPE_OK; ++g) {
        eng_group *G = &E->g[g];
        hype_tile *buf;
        if ((rc = ensure_prime_cap(E, G, count_tiles(E, g, 1))) != HYPE_OK) break;
        buf = G->stage_prime[E->par];
        m = build_tiles(E, g, 1, buf, G->prime_cap);
        if (m < 0) {
            rc = HYPE_E_UNSUPPORTED;
            break;
        }
        G->nprime = m;
        if (m == 0) continue;
        rc = hyped_memcpy_h2d(G->prime_d, buf, (size_t)m * sizeof(hype_tile), E->stream);
        if (rc == HYPE_OK) rc = hyped_fill_u64(G->nprime_d, (unsigned long long)(unsigned)m, 1, E->stream);
        if (rc == HYPE_OK) {
            hype_launch L = base(E, G, HYPE_MODE_PRIME, 0);
            rc = hyped_launch(&L, E->stream);
        }
    }
    for (k = 0; k < E->cap; ++k) E->fresh[k] = 0;
    if (rc == HYPE_OK && !(E->flags & ENG_ASYNC)) rc = hyped_stream_sync(E->stream);
    return rc;
}

static int capture(hype_engine *E)
{
    int g, s, rc, rc2;
    void *graph = NULL;
    /* capture on an idle stream (a new variant or a grown table; rare) */
    HYPE_TRY(hyped_stream_sync(E->stream));
    HYPE_TRY(hyped_graph_begin(E->stream));
    rc = hyped_event_record(E->ev_fork, E->stream);
    for (g = 0; g < E->ng && rc == HYPE_OK; ++g) {
        eng_group *G = &E->g[g];
        rc = hyped_stream_wait(G->stream, E->ev_fork);
        for (s = 0; s < G->nstages && rc == HYPE_OK; ++s) {
            hype_launch L = base(E, G, HYPE_MODE_STAGE, s);
            rc = hyped_launch(&L, G->stream);
        }
        if (rc == HYPE_OK) rc = hyped_event_record(G->ev_done, G->stream);
        if (rc == HYPE_OK) rc = hyped_stream_wait(E->stream, G->ev_done);
    }
    rc2 = hyped_graph_end(E->stream, &graph);
    if (rc != HYPE_OK) return rc;
    if (rc2 != HYPE_OK) return rc2;
    E->graph = graph;
    return HYPE_OK;
}

int eng_step(hype_engine *E, int nsteps)
{
    int k;
    if (E->ng == 0) return HYPE_OK;
    if (!E->graph) HYPE_TRY(capture(E));
    for (k = 0; k < nsteps; ++k) HYPE_TRY(hyped_graph_launch(E->graph, E->stream));
    return HYPE_OK;
}

int eng_sync(hype_engine *E) { return hyped_stream_sync(E->stream); }

int eng_read_states(hype_engine *E, hype_dstate *out)
{
    HYPE_TRY(hyped_memcpy_d2h(out, E->st_d, (size_t)E->cap * sizeof(hype_dstate), E->stream));
    return hyped_stream_sync(E->stream);
}
