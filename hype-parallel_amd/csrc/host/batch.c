/*
 * batch.c -- many independent instances advanced together (host side, C).
 *
 * The reference advances one grid per program run (hype.c:14-456); this
 * driver keeps a set of grids resident in HBM and advances all of them with
 * one persistent launch per SSP-RK stage per kernel variant, replayed from a
 * captured hipGraph.  Every instance keeps its own time state: dt from its
 * own CFL maximum, its own end time and step budget.  The tile height is
 * chosen by the tile-shape model and can be refined by measurement
 * (hype_batch_autotune).
 */
#include <stdlib.h>
#include <string.h>

#include "engine.h"

struct hype_batch {
    hype_engine E;
    int n;
    double *slab;
    size_t *off;      /* per instance, in doubles: U at off, A, B after */
    size_t *ncell;
    int *nv;
    hype_batch_opts opts;
    int began;        /* hype_batch_begin has run: the time states are set */
    void *ev0, *ev1;
    char msg[256];
    double err_x, err_y;
};

void hype_batch_destroy(hype_batch *b)
{
    if (!b) return;
    eng_free(&b->E);
    hyped_free(b->slab);
    if (b->ev0) hyped_event_destroy(b->ev0);
    if (b->ev1) hyped_event_destroy(b->ev1);
    free(b->off);
    free(b->ncell);
    free(b->nv);
    free(b);
}

int hype_batch_create(const hype_problem *probs, int n, const hype_batch_opts *opts, hype_batch **out)
{
    hype_batch *b;
    size_t total = 0;
    int i, rc;
    if (!probs || n < 1 || !out) return HYPE_E_INVALID;
    *out = NULL;
    for (i = 0; i < n; ++i) {
        int nv;
        HYPE_TRY(hype_check_problem(&probs[i], &nv));
        if (probs[i].order < 1) return HYPE_E_UNSUPPORTED;
    }
    b = (hype_batch *)calloc(1, sizeof(*b));
    if (!b) return HYPE_E_NOMEM;
    b->n = n;
    if (opts) b->opts = *opts;
    b->off = (size_t *)calloc((size_t)n, sizeof(size_t));
    b->ncell = (size_t *)calloc((size_t)n, sizeof(size_t));
    b->nv = (int *)calloc((size_t)n, sizeof(int));
    if (!b->off || !b->ncell || !b->nv) {
        hype_batch_destroy(b);
        return HYPE_E_NOMEM;
    }
    for (i = 0; i < n; ++i) {
        b->nv[i] = hype_nvar(probs[i].system);
        b->ncell[i] = (size_t)probs[i].nx * probs[i].ny;
        b->off[i] = total;
        total += 3 * b->ncell[i] * b->nv[i] + hype_inlet_doubles(&probs[i]); /* U, A, B, inlet table */
    }
    if ((rc = eng_init(&b->E, n, b->opts.tile_rows, 0)) != HYPE_OK ||
        (rc = hyped_malloc((void **)&b->slab, total * sizeof(double))) != HYPE_OK ||
        (rc = hyped_event_create(&b->ev0)) != HYPE_OK || (rc = hyped_event_create(&b->ev1)) != HYPE_OK ||
        (rc = hyped_fill_u64(b->slab, 0ull, total, b->E.stream)) != HYPE_OK) {
        hype_batch_destroy(b);
        return rc;
    }
    for (i = 0; i < n; ++i) {
        size_t m = b->ncell[i] * b->nv[i];
        double *U = b->slab + b->off[i], *inlet = NULL;
        const size_t ni = hype_inlet_doubles(&probs[i]);
        if (ni > 0) {
            double *tab = (double *)malloc(ni * sizeof(double));
            if (!tab) rc = HYPE_E_NOMEM;
            else {
                hype_inlet_table(&probs[i], tab);
                inlet = U + 3 * m;
                rc = hyped_memcpy_h2d(inlet, tab, ni * sizeof(double), b->E.stream);
                if (rc == HYPE_OK) rc = hyped_stream_sync(b->E.stream);
                free(tab);
            }
            if (rc != HYPE_OK) {
                hype_batch_destroy(b);
                return rc;
            }
        }
        if ((rc = eng_attach(&b->E, i, &probs[i], U, U + m, U + 2 * m, inlet)) != HYPE_OK) {
            hype_batch_destroy(b);
            return rc;
        }
    }
    if (b->opts.tile_rows < 1) eng_set_rows(&b->E, eng_pick_rows(&b->E));
    if ((rc = eng_commit(&b->E)) != HYPE_OK) {
        hype_batch_destroy(b);
        return rc;
    }
    *out = b;
    return HYPE_OK;
}

int hype_batch_set_state(hype_batch *b, int i, const double *U, int mem)
{
    size_t bytes;
    if (!b || i < 0 || i >= b->n || !U) return HYPE_E_INVALID;
    bytes = b->ncell[i] * b->nv[i] * sizeof(double);
    if (mem == HYPE_MEM_HOST) HYPE_TRY(hyped_memcpy_h2d(b->slab + b->off[i], U, bytes, b->E.stream));
    else HYPE_TRY(hyped_memcpy_d2d(b->slab + b->off[i], U, bytes, b->E.stream));
    return hyped_stream_sync(b->E.stream);
}

int hype_batch_get_state(hype_batch *b, int i, double *U, int mem)
{
    size_t bytes;
    if (!b || i < 0 || i >= b->n || !U) return HYPE_E_INVALID;
    bytes = b->ncell[i] * b->nv[i] * sizeof(double);
    if (mem == HYPE_MEM_HOST) HYPE_TRY(hyped_memcpy_d2h(U, b->slab + b->off[i], bytes, b->E.stream));
    else HYPE_TRY(hyped_memcpy_d2d(U, b->slab + b->off[i], bytes, b->E.stream));
    return hyped_stream_sync(b->E.stream);
}

int hype_batch_begin(hype_batch *b, const double *t0, const double *t_end, const int *max_steps,
                     unsigned flags)
{
    int *slots, *clip, i, rc;
    if (!b || !t_end || !max_steps) return HYPE_E_INVALID;
    slots = (int *)malloc((size_t)b->n * sizeof(int));
    clip = (int *)malloc((size_t)b->n * sizeof(int));
    if (!slots || !clip) {
        free(slots);
        free(clip);
        return HYPE_E_NOMEM;
    }
    for (i = 0; i < b->n; ++i) {
        slots[i] = i;
        clip[i] = (flags & HYPE_TS_CLIP) ? 1 : 0;
    }
    rc = eng_start(&b->E, slots, b->n, t0, t_end, max_steps, clip, NULL, NULL);
    free(slots);
    free(clip);
    if (rc == HYPE_OK) b->began = 1;
    return rc;
}

int hype_batch_advance(hype_batch *b, int nsteps)
{
    if (!b || nsteps < 0 || !b->began) return HYPE_E_INVALID;
    return eng_step(&b->E, nsteps);
}

int hype_batch_advance_timed(hype_batch *b, int nsteps, double *ms)
{
    float f = 0.0f;
    if (!b || nsteps < 0 || !ms || !b->began) return HYPE_E_INVALID;
    HYPE_TRY(hyped_event_record(b->ev0, b->E.stream));
    HYPE_TRY(eng_step(&b->E, nsteps));
    HYPE_TRY(hyped_event_record(b->ev1, b->E.stream));
    HYPE_TRY(hyped_event_elapsed_ms(b->ev0, b->ev1, &f));
    *ms = (double)f;
    return HYPE_OK;
}

int hype_batch_sync(hype_batch *b, hype_status *st)
{
    hype_dstate *d;
    int i, rc = HYPE_OK;
    if (!b) return HYPE_E_INVALID;
    d = (hype_dstate *)malloc((size_t)b->n * sizeof(hype_dstate));
    if (!d) return HYPE_E_NOMEM;
    rc = eng_read_states(&b->E, d);
    for (i = 0; rc == HYPE_OK && i < b->n; ++i) {
        unsigned long long e = d[i].step > 0 ? d[i].err[(d[i].step - 1) & 1] : d[i].err[1];
        if (st) {
            st[i].t = d[i].t;
            st[i].dt = d[i].dt;
            st[i].steps = d[i].step;
            st[i].done = d[i].done || d[i].step >= d[i].max_steps || !(d[i].t < d[i].t_end) || e != ~0ull;
            st[i].err = e != ~0ull ? HYPE_E_INADMISSIBLE : HYPE_OK;
        }
        if (e != ~0ull && rc == HYPE_OK) {
            const hype_problem *p = &b->E.P[i];
            hype_decode_error(e, p->nx, p->ny, p->order < 2 ? 1 : 2, p->x_min, p->y_min,
                              (p->x_max - p->x_min) / (double)p->nx, b->msg, sizeof(b->msg), &b->err_x,
                              &b->err_y);
            rc = HYPE_E_INADMISSIBLE;
        }
    }
    free(d);
    return rc;
}

const char *hype_batch_last_error(hype_batch *b, double *x, double *y)
{
    if (!b) return "no batch";
    if (x) *x = b->err_x;
    if (y) *y = b->err_y;
    return b->msg;
}

int hype_batch_tiling(hype_batch *b, int *rows, int *tiles, int *groups)
{
    if (!b) return HYPE_E_INVALID;
    if (rows) *rows = b->E.target_rows;
    if (tiles) *tiles = eng_total_tiles(&b->E);
    if (groups) *groups = b->E.ng;
    return HYPE_OK;
}

int hype_batch_set_rows(hype_batch *b, int rows)
{
    if (!b) return HYPE_E_INVALID;
    HYPE_TRY(eng_set_rows(&b->E, rows));
    return eng_commit(&b->E);
}

/* Measured tuning of the tile shape and LDS footprint: time `steps` steps
 * of every candidate tile height, and, where an instance's system is built
 * for both tile layouts (one wave per tile with one LDS window, or tiles
 * split over two waves with two windows and the exchange areas), of both
 * layouts; keep the fastest pair.  The states are advanced from the current
 * ones: set them again afterwards. */
int hype_batch_autotune(hype_batch *b, const int *cand, int ncand, int steps, int *best_rows,
                        double *best_ms)
{
    int k, i, l, best = -1, best_layout = -1, nlay = 1, rc = HYPE_OK;
    int layouts[2] = {-1, 1};
    double bms = 1e300;
    double *t_end;
    int *maxs;
    if (!b || !cand || ncand < 1 || steps < 1) return HYPE_E_INVALID;
    for (i = 0; i < b->n; ++i)
        if (hype_layouts(&b->E.P[i], hyped_pair_modes(b->E.P[i].system)) == 3) nlay = 2;
    t_end = (double *)malloc((size_t)b->n * sizeof(double));
    maxs = (int *)malloc((size_t)b->n * sizeof(int));
    if (!t_end || !maxs) {
        free(t_end);
        free(maxs);
        return HYPE_E_NOMEM;
    }
    for (i = 0; i < b->n; ++i) {
        t_end[i] = 1e300;
        maxs[i] = 1 << 30;
    }
    for (l = 0; l < nlay && rc == HYPE_OK; ++l) {
        if ((rc = eng_set_layout(&b->E, layouts[l])) != HYPE_OK) break;
        for (k = 0; k < ncand; ++k) {
            double ms = 0.0;
            if (cand[k] < 1) continue;
            if ((rc = hype_batch_set_rows(b, cand[k])) != HYPE_OK ||
                (rc = hype_batch_begin(b, NULL, t_end, maxs, 0)) != HYPE_OK ||
                (rc = hype_batch_advance(b, 1)) != HYPE_OK ||
                (rc = hype_batch_advance_timed(b, steps, &ms)) != HYPE_OK)
                break;
            if (ms < bms) {
                bms = ms;
                best = cand[k];
                best_layout = layouts[l];
            }
        }
    }
    free(t_end);
    free(maxs);
    if (rc != HYPE_OK) return rc;
    if (best < 1) return HYPE_E_INVALID;
    if (best_rows) *best_rows = best;
    if (best_ms) *best_ms = bms;
    HYPE_TRY(eng_set_layout(&b->E, best_layout));
    return hype_batch_set_rows(b, best);
}

int hype_batch_layout(hype_batch *b, int *layout)
{
    if (!b || !layout) return HYPE_E_INVALID;
    *layout = b->E.layout;
    return HYPE_OK;
}
