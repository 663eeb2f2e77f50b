/*
 * ctx.c -- single-instance context: the RHSFunction drop-in and the
 * device-resident SSP-RK driver (host side, C).
 *
 * Reference counterparts: RHSFunctionPrim / RHSFunction (rhs_function_prim.c:12,
 * rhs_function.c:12), the TS setup and solve of hype.c:337-384, and the dt
 * hand-off of MonitorFunction (post_process.c:19).  The host never sees dt
 * during a solve: the last stage of every step reduces the wave speed on the
 * device and the first stage of the next step turns it into dt.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

struct hype_ctx {
    hype_problem P;
    int nv, nstages, nsy, ntiles, grid, slots;
    int pair; /* 1: split tiles on two-wave workgroups */
    size_t ncell, nbytes;
    void *stream;
    double *buf[3];         /* U, A, B (device) */
    double *R;              /* RHS output (device) */
    hype_dstate *st;        /* device time state */
    hype_inst *inst_d;
    hype_tile *tiles_d;
    unsigned *queue_d;
    double *hist_d;
    int hist_cap;
    unsigned char *fb_d;
    double *inlet_d;        /* left-inlet table (device), or NULL */
    hype_inst inst_h;
    void *graph;            /* one captured step */
    unsigned flags;
    int began;
    char msg[256];
    double err_x, err_y;
};

static void ctx_free(hype_ctx *c)
{
    int i;
    if (!c) return;
    if (c->graph) hyped_graph_destroy(c->graph);
    for (i = 0; i < 3; ++i) hyped_free(c->buf[i]);
    hyped_free(c->R);
    hyped_free(c->st);
    hyped_free(c->inst_d);
    hyped_free(c->tiles_d);
    hyped_free(c->queue_d);
    hyped_free(c->hist_d);
    hyped_free(c->fb_d);
    hyped_free(c->inlet_d);
    if (c->stream) hyped_stream_destroy(c->stream);
    free(c);
}

int hype_ctx_create(const hype_problem *p, hype_ctx **out)
{
    hype_ctx *c;
    hype_tile *tiles;
    int rc, nv = 0, i;
    if (!out) return HYPE_E_INVALID;
    *out = NULL;
    HYPE_TRY(hype_check_problem(p, &nv));
    HYPE_TRY(hyped_init());
    c = (hype_ctx *)calloc(1, sizeof(*c));
    if (!c) return HYPE_E_NOMEM;
    c->P = *p;
    c->nv = nv;
    c->nstages = hyped_num_stages(p->order);
    c->ncell = (size_t)p->nx * p->ny;
    c->nbytes = c->ncell * nv * sizeof(double);
    /* tile layout and row split: the tile-shape model's cheapest among the
     * layouts the system's kernel is built for */
    if (hype_pick_layout(p, &c->pair, &c->nsy, &c->slots) != HYPE_OK) {
        free(c);
        return HYPE_E_UNSUPPORTED;
    }
    c->ntiles = hype_strips(p->nx) * c->nsy;
    c->grid = 0;
    if ((rc = hyped_stream_create(&c->stream)) != HYPE_OK) goto fail;
    for (i = 0; i < 3; ++i)
        if ((rc = hyped_malloc((void **)&c->buf[i], c->nbytes)) != HYPE_OK) goto fail;
    if ((rc = hyped_malloc((void **)&c->R, c->nbytes)) != HYPE_OK) goto fail;
    if ((rc = hyped_malloc((void **)&c->st, sizeof(hype_dstate))) != HYPE_OK) goto fail;
    if ((rc = hyped_malloc((void **)&c->inst_d, sizeof(hype_inst))) != HYPE_OK) goto fail;
    if ((rc = hyped_malloc((void **)&c->tiles_d, (size_t)c->ntiles * sizeof(hype_tile))) != HYPE_OK) goto fail;
    if ((rc = hyped_malloc((void **)&c->queue_d, 4 * sizeof(unsigned))) != HYPE_OK) goto fail;
    if ((rc = hyped_fill_u64(c->queue_d, 0ull, 2, c->stream)) != HYPE_OK) goto fail;

    hype_fill_inst(&c->inst_h, p, c->nsy);
    c->inst_h.buf[0] = c->buf[0];
    c->inst_h.buf[1] = c->buf[1];
    c->inst_h.buf[2] = c->buf[2];
    c->inst_h.buf[3] = NULL;
    c->inst_h.st = c->st;
    if (hype_inlet_doubles(p) > 0) {
        const size_t n = hype_inlet_doubles(p);
        double *tab = (double *)malloc(n * sizeof(double));
        if (!tab) { rc = HYPE_E_NOMEM; goto fail; }
        hype_inlet_table(p, tab);
        rc = hyped_malloc((void **)&c->inlet_d, n * sizeof(double));
        if (rc == HYPE_OK) rc = hyped_memcpy_h2d(c->inlet_d, tab, n * sizeof(double), c->stream);
        if (rc == HYPE_OK) rc = hyped_stream_sync(c->stream);
        free(tab);
        if (rc != HYPE_OK) goto fail;
        c->inst_h.inlet = c->inlet_d;
    }
    if ((rc = hyped_memcpy_h2d(c->inst_d, &c->inst_h, sizeof(hype_inst), c->stream)) != HYPE_OK) goto fail;
    tiles = (hype_tile *)malloc((size_t)c->ntiles * sizeof(hype_tile));
    if (!tiles) { rc = HYPE_E_NOMEM; goto fail; }
    if (hype_make_tiles(0, p->nx, p->ny, hype_strips(p->nx), c->nsy, tiles, c->ntiles) != c->ntiles)
        rc = HYPE_E_INVALID;
    else if ((rc = hype_xcd_order(tiles, c->ntiles, HYPE_NXCD)) == HYPE_OK)
        rc = hyped_memcpy_h2d(c->tiles_d, tiles, (size_t)c->ntiles * sizeof(hype_tile), c->stream);
    if (rc == HYPE_OK) rc = hyped_stream_sync(c->stream);
    free(tiles);
    if (rc != HYPE_OK) goto fail;
    *out = c;
    return HYPE_OK;
fail:
    ctx_free(c);
    return rc;
}

void hype_ctx_destroy(hype_ctx *c) { ctx_free(c); }

void *hype_stream(hype_ctx *c) { return c ? c->stream : NULL; }

static hype_launch base_launch(hype_ctx *c, int mode, int stage)
{
    hype_launch L;
    memset(&L, 0, sizeof(L));
    L.v.system = c->P.system;
    L.v.prim = c->P.recons_prim;
    L.v.order = c->P.order;
    L.v.fma = (c->P.options & HYPE_OPT_FMA) != 0;
    L.v.pair = c->pair;
    L.v.inlet = c->P.bc[0] == HYPE_BC_INLET;
    L.insts = c->inst_d;
    L.tiles = c->tiles_d;
    L.ntiles = c->ntiles;
    L.mode = mode;
    L.stage = stage;
    L.grid = c->grid;
    L.queue = c->queue_d;
    return L;
}

static int set_error(hype_ctx *c, unsigned long long word)
{
    int ngf = c->P.order < 2 ? 1 : 2;
    double h = (c->P.x_max - c->P.x_min) / (double)c->P.nx;
    hype_decode_error(word, c->P.nx, c->P.ny, ngf, c->P.x_min, c->P.y_min, h, c->msg,
                      sizeof(c->msg), &c->err_x, &c->err_y);
    return HYPE_E_INADMISSIBLE;
}

const char *hype_last_error(hype_ctx *c, double *x, double *y)
{
    if (!c) return "no context";
    if (x) *x = c->err_x;
    if (y) *y = c->err_y;
    return c->msg;
}

static int rhs_impl(hype_ctx *c, const double *U, double *RHS, double *dt_out,
                    unsigned char *fallback, int mem)
{
    unsigned long long slots[2] = {0ull, ~0ull}; /* smax_rhs, err_rhs */
    const double *in;
    double *out;
    hype_launch L;
    if (!c || !U || !RHS) return HYPE_E_INVALID;
    if (mem == HYPE_MEM_HOST) {
        HYPE_TRY(hyped_memcpy_h2d(c->buf[1], U, c->nbytes, c->stream));
        in = c->buf[1];
        out = c->R;
    } else {
        in = U;
        out = RHS;
    }
    HYPE_TRY(hyped_memcpy_h2d(&c->st->smax_rhs, slots, sizeof(slots), c->stream));
    L = base_launch(c, HYPE_MODE_RHS, 0);
    L.rhs_in = in;
    L.rhs_out = out;
    if (fallback) {
        if (!c->fb_d) HYPE_TRY(hyped_malloc((void **)&c->fb_d, c->ncell));
        L.fallback = c->fb_d;
    }
    HYPE_TRY(hyped_launch(&L, c->stream));
    if (mem == HYPE_MEM_HOST) HYPE_TRY(hyped_memcpy_d2h(RHS, c->R, c->nbytes, c->stream));
    if (fallback) HYPE_TRY(hyped_memcpy_d2h(fallback, c->fb_d, c->ncell, c->stream));
    HYPE_TRY(hyped_memcpy_d2h(slots, &c->st->smax_rhs, sizeof(slots), c->stream));
    HYPE_TRY(hyped_stream_sync(c->stream));
    if (dt_out) {
        double smax, h = (c->P.x_max - c->P.x_min) / (double)(c->P.nx);
        memcpy(&smax, &slots[0], sizeof(smax));
        *dt_out = (c->P.cfl * h) / (2.0 * smax); /* rhs_function_prim.c:620 */
    }
    if (slots[1] != ~0ull) return set_error(c, slots[1]);
    return HYPE_OK;
}

int hype_rhs(hype_ctx *c, double t, const double *U, double *RHS, double *dt_out, int mem)
{
    (void)t;
    return rhs_impl(c, U, RHS, dt_out, NULL, mem);
}

int hype_rhs_diag(hype_ctx *c, double t, const double *U, double *RHS, double *dt_out,
                  unsigned char *fallback, int mem)
{
    (void)t;
    return rhs_impl(c, U, RHS, dt_out, fallback, mem);
}

int hype_cons2prim(hype_ctx *c, const double *U, double *W, int mem)
{
    if (!c || !U || !W) return HYPE_E_INVALID;
    if (mem == HYPE_MEM_HOST) {
        HYPE_TRY(hyped_memcpy_h2d(c->buf[1], U, c->nbytes, c->stream));
        HYPE_TRY(hyped_cons2prim(c->P.system, c->buf[1], c->R, (long)c->ncell, c->P.gamma1,
                                 c->P.gamma2, c->P.pi1, c->P.pi2, c->stream));
        HYPE_TRY(hyped_memcpy_d2h(W, c->R, c->nbytes, c->stream));
    } else {
        HYPE_TRY(hyped_cons2prim(c->P.system, U, W, (long)c->ncell, c->P.gamma1, c->P.gamma2,
                                 c->P.pi1, c->P.pi2, c->stream));
    }
    return hyped_stream_sync(c->stream);
}

int hype_set_state(hype_ctx *c, const double *U, int mem)
{
    if (!c || !U) return HYPE_E_INVALID;
    if (mem == HYPE_MEM_HOST) HYPE_TRY(hyped_memcpy_h2d(c->buf[0], U, c->nbytes, c->stream));
    else HYPE_TRY(hyped_memcpy_d2d(c->buf[0], U, c->nbytes, c->stream));
    return hyped_stream_sync(c->stream);
}

int hype_get_state(hype_ctx *c, double *U, int mem)
{
    if (!c || !U) return HYPE_E_INVALID;
    if (mem == HYPE_MEM_HOST) HYPE_TRY(hyped_memcpy_d2h(U, c->buf[0], c->nbytes, c->stream));
    else HYPE_TRY(hyped_memcpy_d2d(U, c->buf[0], c->nbytes, c->stream));
    return hyped_stream_sync(c->stream);
}

/* the launches of one time step, in order */
static int enqueue_step(hype_ctx *c)
{
    int s;
    for (s = 0; s < c->nstages; ++s) {
        hype_launch L = base_launch(c, HYPE_MODE_STAGE, s);
        HYPE_TRY(hyped_launch(&L, c->stream));
    }
    return HYPE_OK;
}

int hype_begin(hype_ctx *c, double t0, double t_end, int max_steps, unsigned flags)
{
    hype_dstate d;
    hype_launch L;
    if (!c || max_steps < 0) return HYPE_E_INVALID;
    if (c->P.order == 0) return HYPE_E_UNSUPPORTED; /* forward Euler stepping not on the device */
    if (max_steps > c->hist_cap) {
        hyped_free(c->hist_d);
        c->hist_d = NULL;
        c->hist_cap = 0;
        if (max_steps > 0) {
            HYPE_TRY(hyped_malloc((void **)&c->hist_d, (size_t)max_steps * sizeof(double)));
            c->hist_cap = max_steps;
        }
    }
    memset(&d, 0, sizeof(d));
    d.t = t0;
    d.dt = 0.0;
    d.t_end = t_end;
    d.smax[0] = 0ull;
    d.smax[1] = 0ull;
    d.smax_rhs = 0ull;
    d.err[0] = ~0ull;
    d.err[1] = ~0ull;
    d.err_rhs = ~0ull;
    d.step = 0;
    d.cur = 0;
    d.max_steps = max_steps;
    d.done = 0;
    d.clip = (flags & HYPE_TS_CLIP) ? 1 : 0;
    d.hist_len = c->hist_cap;
    d.dt_hist = c->hist_d;
    HYPE_TRY(hyped_memcpy_h2d(c->st, &d, sizeof(d), c->stream));
    /* d lives on this stack frame: the copy completes before it goes */
    HYPE_TRY(hyped_stream_sync(c->stream));
    /* priming call (hype.c:343): the wave-speed maximum of U^0 */
    L = base_launch(c, HYPE_MODE_PRIME, 0);
    HYPE_TRY(hyped_launch(&L, c->stream));
    c->flags = flags;
    c->began = 1;
    if ((flags & HYPE_TS_GRAPH) && !c->graph) {
        HYPE_TRY(hyped_graph_begin(c->stream));
        {
            int rc = enqueue_step(c);
            void *g = NULL;
            int rc2 = hyped_graph_end(c->stream, &g);
            if (rc != HYPE_OK) return rc;
            if (rc2 != HYPE_OK) return rc2;
            c->graph = g;
        }
    }
    return HYPE_OK;
}

int hype_advance(hype_ctx *c, int nsteps)
{
    int k;
    if (!c || !c->began) return HYPE_E_INVALID;
    for (k = 0; k < nsteps; ++k) {
        if ((c->flags & HYPE_TS_GRAPH) && c->graph) HYPE_TRY(hyped_graph_launch(c->graph, c->stream));
        else HYPE_TRY(enqueue_step(c));
    }
    return HYPE_OK;
}

int hype_advance_timed(hype_ctx *c, int nsteps, double *ms)
{
    void *e0 = NULL, *e1 = NULL;
    float f = 0.0f;
    int rc;
    if (!c || !ms) return HYPE_E_INVALID;
    HYPE_TRY(hyped_event_create(&e0));
    rc = hyped_event_create(&e1);
    if (rc == HYPE_OK) rc = hyped_event_record(e0, c->stream);
    if (rc == HYPE_OK) rc = hype_advance(c, nsteps);
    if (rc == HYPE_OK) rc = hyped_event_record(e1, c->stream);
    if (rc == HYPE_OK) rc = hyped_event_elapsed_ms(e0, e1, &f);
    *ms = (double)f;
    hyped_event_destroy(e0);
    hyped_event_destroy(e1);
    return rc;
}

int hype_sync(hype_ctx *c, hype_status *st)
{
    hype_dstate d;
    if (!c) return HYPE_E_INVALID;
    unsigned long long e;
    HYPE_TRY(hyped_memcpy_d2h(&d, c->st, sizeof(d), c->stream));
    HYPE_TRY(hyped_stream_sync(c->stream));
    /* the failing step is the last one taken (stepping stops after it);
     * with no step taken, a failure belongs to the priming call */
    e = d.step > 0 ? d.err[(d.step - 1) & 1] : d.err[1];
    if (st) {
        st->t = d.t;
        st->dt = d.dt;
        st->steps = d.step;
        st->done = d.done || d.step >= d.max_steps || !(d.t < d.t_end) || e != ~0ull;
        st->err = (e != ~0ull) ? HYPE_E_INADMISSIBLE : HYPE_OK;
    }
    if (e != ~0ull) return set_error(c, e);
    return HYPE_OK;
}

int hype_dt_history(hype_ctx *c, double *dts, int n)
{
    if (!c || !dts || n < 0) return HYPE_E_INVALID;
    if (n > c->hist_cap) n = c->hist_cap;
    if (n == 0) return HYPE_OK;
    HYPE_TRY(hyped_memcpy_d2h(dts, c->hist_d, (size_t)n * sizeof(double), c->stream));
    return hyped_stream_sync(c->stream);
}

/* tiling hooks: rows per tile (the grid is split into ceil(ny/hrows)
 * near-equal segments) and the persistent grid size */
int hype_ctx_set_tiling(hype_ctx *c, int hrows, int grid)
{
    hype_tile *tiles, *tiles_d = NULL;
    hype_inst inst;
    int rc, nsy, ntiles;
    if (!c || hrows < 1 || grid < 0) return HYPE_E_INVALID;
    nsy = (c->P.ny + hrows - 1) / hrows;
    ntiles = hype_strips(c->P.nx) * nsy;
    /* the new table is built and uploaded first; the context switches to it
     * only when every step succeeded, and keeps the old one otherwise */
    tiles = (hype_tile *)malloc((size_t)ntiles * sizeof(hype_tile));
    if (!tiles) return HYPE_E_NOMEM;
    if (hype_make_tiles(0, c->P.nx, c->P.ny, hype_strips(c->P.nx), nsy, tiles, ntiles) != ntiles) {
        free(tiles);
        return HYPE_E_INVALID;
    }
    rc = hype_xcd_order(tiles, ntiles, HYPE_NXCD);
    if (rc == HYPE_OK) rc = hyped_malloc((void **)&tiles_d, (size_t)ntiles * sizeof(hype_tile));
    if (rc == HYPE_OK) rc = hyped_memcpy_h2d(tiles_d, tiles, (size_t)ntiles * sizeof(hype_tile), c->stream);
    inst = c->inst_h;
    inst.nsy = nsy;
    if (rc == HYPE_OK) rc = hyped_memcpy_h2d(c->inst_d, &inst, sizeof(hype_inst), c->stream);
    if (rc == HYPE_OK) rc = hyped_stream_sync(c->stream);
    free(tiles);
    if (rc != HYPE_OK) {
        hyped_free(tiles_d);
        return rc;
    }
    if (c->graph) {
        hyped_graph_destroy(c->graph);
        c->graph = NULL;
    }
    hyped_free(c->tiles_d);
    c->tiles_d = tiles_d;
    c->inst_h = inst;
    c->nsy = nsy;
    c->grid = grid;
    c->ntiles = ntiles;
    return HYPE_OK;
}

int hype_ctx_tiling(hype_ctx *c, int *hrows, int *ntiles, int *grid)
{
    if (!c) return HYPE_E_INVALID;
    if (hrows) *hrows = (c->P.ny + c->nsy - 1) / c->nsy;
    if (ntiles) *ntiles = c->ntiles;
    if (grid) *grid = c->grid;
    return HYPE_OK;
}
