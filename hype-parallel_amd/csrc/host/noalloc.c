/*
 * noalloc.c -- the RHS callback without allocation (host side, C).
 *
 * hype_rhs_noalloc is hype_rhs on caller-owned device buffers for callers
 * that cannot allocate: it creates no context, calls no device or host
 * allocator, and keeps what a context would hold in device memory (instance
 * descriptor, time state, tile table, left-inlet table) in the library's
 * fixed device scratch (scratch.hip, HYPE_SCRATCH_BYTES per GPU).  The tile
 * layout is the one hype_ctx_create picks, so the kernels and their results
 * are those of hype_rhs.
 *
 * The scratch and the host staging below are shared by every call in the
 * process; a mutex serialises the calls, and each call waits for its stream
 * before it returns, so no two calls use the scratch at once.
 */
#include <pthread.h>
#include <string.h>

#include "common.h"

static pthread_mutex_t na_mu = PTHREAD_MUTEX_INITIALIZER;
/* host staging of one call: descriptor and time state, tile table (model
 * order, then XCD order), inlet table */
static unsigned char na_head[HYPE_SCRATCH_TILES_OFF];
static hype_tile na_rows[HYPE_SCRATCH_MAX_TILES];
static hype_tile na_tiles[HYPE_SCRATCH_MAX_TILES];
static double na_inlet[HYPE_SCRATCH_INLET_DOUBLES];

static int rhs_noalloc_locked(const hype_problem *p, const double *U, double *RHS, double *dt_out,
                              double *err_xy, void *stream)
{
    unsigned char *base = NULL;
    size_t bytes = 0, ninlet;
    hype_inst inst;
    hype_dstate st;
    hype_launch L;
    unsigned long long slots[2] = {0ull, ~0ull}; /* smax_rhs, err_rhs */
    int nv = 0, pair = 0, nsy = 1, nsx, ntiles, rc, sync_rc;
    HYPE_TRY(hype_check_problem(p, &nv));
    HYPE_TRY(hyped_init());
    HYPE_TRY(hyped_scratch((void **)&base, &bytes));
    if (bytes < HYPE_SCRATCH_BYTES) return HYPE_E_DEVICE;
    /* sizes the scratch holds: the tile table and the inlet table */
    nsx = hype_strips(p->nx);
    if (nsx > HYPE_SCRATCH_MAX_TILES) return HYPE_E_TOO_LARGE;
    ninlet = hype_inlet_doubles(p);
    if (ninlet > HYPE_SCRATCH_INLET_DOUBLES) return HYPE_E_TOO_LARGE;
    HYPE_TRY(hype_pick_layout(p, &pair, &nsy, NULL));
    if ((long)nsx * nsy > HYPE_SCRATCH_MAX_TILES) nsy = HYPE_SCRATCH_MAX_TILES / nsx; /* taller tiles */
    ntiles = nsx * nsy;
    if (hype_make_tiles(0, p->nx, p->ny, nsx, nsy, na_rows, HYPE_SCRATCH_MAX_TILES) != ntiles)
        return HYPE_E_INVALID;
    hype_xcd_order_into(na_rows, na_tiles, ntiles, HYPE_NXCD);

    hype_fill_inst(&inst, p, nsy);
    inst.buf[0] = (double *)U;
    inst.buf[1] = RHS;
    inst.st = (hype_dstate *)(base + HYPE_SCRATCH_STATE_OFF);
    if (ninlet > 0) {
        hype_inlet_table(p, na_inlet);
        inst.inlet = (const double *)(base + HYPE_SCRATCH_INLET_OFF);
    }
    /* the kernel's RHS mode reduces into smax_rhs and err_rhs only */
    memset(&st, 0, sizeof(st));
    st.smax_rhs = 0ull;
    st.err_rhs = ~0ull;
    st.err[0] = st.err[1] = ~0ull;
    memset(na_head, 0, sizeof(na_head));
    memcpy(na_head, &inst, sizeof(inst));
    memcpy(na_head + HYPE_SCRATCH_STATE_OFF, &st, sizeof(st));

    /* from here on a failure waits for the stream before it returns, so
     * no copy out of the static staging is still pending for the next call */
    rc = hyped_memcpy_h2d(base, na_head, sizeof(na_head), stream);
    if (rc == HYPE_OK)
        rc = hyped_memcpy_h2d(base + HYPE_SCRATCH_TILES_OFF, na_tiles, (size_t)ntiles * sizeof(hype_tile), stream);
    if (rc == HYPE_OK && ninlet > 0)
        rc = hyped_memcpy_h2d(base + HYPE_SCRATCH_INLET_OFF, na_inlet, ninlet * sizeof(double), stream);

    memset(&L, 0, sizeof(L));
    L.v.system = p->system;
    L.v.prim = p->recons_prim;
    L.v.order = p->order;
    L.v.fma = (p->options & HYPE_OPT_FMA) != 0;
    L.v.pair = pair;
    L.v.inlet = ninlet > 0;
    L.insts = (const hype_inst *)base;
    L.tiles = (const hype_tile *)(base + HYPE_SCRATCH_TILES_OFF);
    L.ntiles = ntiles;
    L.mode = HYPE_MODE_RHS;
    L.grid = 0; /* one workgroup per tile: no queue counters */
    L.rhs_in = U;
    L.rhs_out = RHS;
    if (rc == HYPE_OK) rc = hyped_launch(&L, stream);
    if (rc == HYPE_OK)
        rc = hyped_memcpy_d2h(slots, base + HYPE_SCRATCH_STATE_OFF + offsetof(hype_dstate, smax_rhs),
                              sizeof(slots), stream);
    sync_rc = hyped_stream_sync(stream);
    if (rc != HYPE_OK) return rc;
    if (sync_rc != HYPE_OK) return sync_rc;
    if (dt_out) {
        double smax, h = (p->x_max - p->x_min) / (double)(p->nx);
        memcpy(&smax, &slots[0], sizeof(smax));
        *dt_out = (p->cfl * h) / (2.0 * smax); /* rhs_function_prim.c:620 */
    }
    if (slots[1] != ~0ull) {
        char msg[256];
        double x = 0.0, y = 0.0;
        hype_decode_error(slots[1], p->nx, p->ny, p->order < 2 ? 1 : 2, p->x_min, p->y_min,
                          (p->x_max - p->x_min) / (double)p->nx, msg, sizeof(msg), &x, &y);
        if (err_xy) {
            err_xy[0] = x;
            err_xy[1] = y;
        }
        return HYPE_E_INADMISSIBLE;
    }
    return HYPE_OK;
}

int hype_rhs_noalloc(const hype_problem *p, double t, const double *U, double *RHS, double *dt_out,
                     double *err_xy, void *stream)
{
    int rc;
    (void)t;
    if (!p || !U || !RHS) return HYPE_E_INVALID;
    pthread_mutex_lock(&na_mu);
    rc = rhs_noalloc_locked(p, U, RHS, dt_out, err_xy, stream);
    pthread_mutex_unlock(&na_mu);
    return rc;
}
