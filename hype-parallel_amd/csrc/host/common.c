/* common.c -- problem validation, tiling and error decoding (host side). */
#include "common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int hype_nvar(int system)
{
    switch (system) {
    case HYPE_EULER: return 4;
    case HYPE_KAPILA: return 6;
    case HYPE_BAER_NUNZIATO: return 9;
    default: return 0;
    }
}

int hype_check_problem(const hype_problem *p, int *nvar)
{
    int s, nv;
    if (!p) return HYPE_E_INVALID;
    nv = hype_nvar(p->system);
    if (nv == 0) return HYPE_E_INVALID;
    if (p->order < 0 || p->order > 3) return HYPE_E_UNSUPPORTED;
    if (p->nx < 3 || p->ny < 3) return HYPE_E_INVALID;
    /* face keys of the error word hold at most 36 bits (hype_key_bits) */
    if (((double)(p->nx + 1) * p->ny + (double)p->nx * (p->ny + 1)) * 2.0 >= 68719476736.0)
        return HYPE_E_UNSUPPORTED;
    if (!(p->x_max > p->x_min) || !(p->y_max > p->y_min)) return HYPE_E_INVALID;
    if (!(p->cfl > 0.0)) return HYPE_E_INVALID;
    if (p->options & ~(HYPE_OPT_FMA | HYPE_OPT_ONE_WAVE_TILES | HYPE_OPT_SPLIT_TILES)) return HYPE_E_INVALID;
    if ((p->options & HYPE_OPT_ONE_WAVE_TILES) && (p->options & HYPE_OPT_SPLIT_TILES)) return HYPE_E_INVALID;
    for (s = 0; s < 4; ++s) {
        if (p->bc[s] == HYPE_BC_WALL) return HYPE_E_UNSUPPORTED;
        /* the reference has an inlet on the left boundary of the Kapila
         * system only (rhs_function_prim.c:135-144; InletBC is defined in
         * kapila_5eqn.c alone) */
        if (p->bc[s] == HYPE_BC_INLET && (s != 0 || p->system != HYPE_KAPILA)) return HYPE_E_UNSUPPORTED;
        if (p->bc[s] != HYPE_BC_PERIODIC && p->bc[s] != HYPE_BC_REFLECTIVE &&
            p->bc[s] != HYPE_BC_TRANSMISSIVE && p->bc[s] != HYPE_BC_INLET)
            return HYPE_E_INVALID;
    }
    if (nvar) *nvar = nv;
    return HYPE_OK;
}

/* near-equal split of n into k parts: part i is [i*n/k, (i+1)*n/k) */
static int part_lo(int i, int n, int k) { return (int)((long)i * n / k); }

int hype_make_tiles(int inst, int nx, int ny, int nsx, int nsy, hype_tile *tiles, int cap)
{
    int n = 0, a, b;
    if (nsx < 1 || nsy < 1 || nsx > nx || nsy > ny) return -1;
    for (b = 0; b < nsy; ++b)
        for (a = 0; a < nsx; ++a) {
            if (n >= cap) return -1;
            tiles[n].inst = inst;
            tiles[n].x0 = part_lo(a, nx, nsx);
            tiles[n].w = part_lo(a + 1, nx, nsx) - tiles[n].x0;
            tiles[n].y0 = part_lo(b, ny, nsy);
            tiles[n].h = part_lo(b + 1, ny, nsy) - tiles[n].y0;
            tiles[n].flags = (n == 0) ? 1 : 0;
            if (tiles[n].w > HYPE_TILE_W || tiles[n].w < 1 || tiles[n].h < 1) return -1;
            ++n;
        }
    return n;
}

int hype_strips(int nx) { return (nx + HYPE_TILE_W - 1) / HYPE_TILE_W; }

/* Tile-shape model.  A wave sweeps its strip one row per iteration with all
 * 64 lanes busy whatever the strip width, so a tile of h owned rows costs
 * about (h + 2 + c0) row-iterations (two ring rows, c0 for the 5-row
 * prologue).  Where the two halves of a tile run on two waves (pair), each
 * wave sweeps about h/2 rows plus one ring row, and the two meet at one
 * face (a quarter of an iteration).  The kernel is bound by FP64 issue:
 * waves that share a SIMD share its issue slots, gaining only latency
 * hiding (a measured 1.5x for two waves).  With w = ceil(waves / simds)
 * waves per SIMD and `occ` resident waves per SIMD, a launch therefore takes
 * about
 *   per * w / gain(w)                      if w <= occ,
 *   per * ceil(w / occ) * occ / gain(occ)   otherwise,
 * with per the row-iterations of one wave. */
static double occ_gain(long w) { return w >= 2 ? 1.5 : 1.0; }

double hype_tile_rows(int h, int pair)
{
    const double c0 = 1.0;
    if (pair && h >= 2) return (double)((h + 1) / 2) + 1.0 + c0 + 0.25;
    return h + 2.0 + c0;
}

double hype_tile_cost(long waves, double per, int simds, int occ)
{
    long w;
    if (simds < 1) simds = 1;
    if (occ < 1) occ = 1;
    w = (waves + simds - 1) / simds;
    if (w <= occ) return per * (double)w / occ_gain(w);
    return per * (double)((w + occ - 1) / occ) * occ / occ_gain(occ);
}

int hype_pick_segments(int nx, int ny, int simds, int occ, int extra_waves, int pair)
{
    int nsx = hype_strips(nx), k, best = 1;
    double best_cost = 1e300;
    for (k = 1; k <= ny; ++k) {
        long waves = (long)nsx * k * (pair ? 2 : 1) + extra_waves;
        double cost = hype_tile_cost(waves, hype_tile_rows((ny + k - 1) / k, pair), simds, occ);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = k;
        }
    }
    return best;
}

/* face-key bits of the error word: every face point of the grid
 * ((nx+1)*ny x-face and nx*(ny+1) y-face points, 2 per face) must fit */
int hype_key_bits(int nx, int ny)
{
    const double faces = ((double)(nx + 1) * ny + (double)nx * (ny + 1)) * 2.0;
    return faces < 268435456.0 ? HYPE_ERR_KEYBITS_SMALL : HYPE_ERR_KEYBITS_LARGE;
}

void hype_decode_error(unsigned long long word, int nx, int ny, int ngf,
                       double x_min, double y_min, double h, char *msg,
                       size_t cap, double *x, double *y)
{
    static const char *what[] = {
        "", "Negative density = %f", "Negative pressure, p = %f",
        "Negative square speed of sound = %f", "Negative density (solid) = %f",
        "Negative density (gas) = %f", "Negative pressure (solid), p = %f",
        "Negative pressure (gas), p = %f"};
    const int K = hype_key_bits(nx, ny);
    long key = (long)((word >> (HYPE_ERR_SEQ_SHIFT - K)) & ((1ull << K) - 1ull));
    int kind = (int)((word >> (HYPE_ERR_SEQ_SHIFT - K - 4)) & 0xF);
    unsigned vb = (unsigned)((word & ((1ull << (56 - K)) - 1ull)) << (K - 24));
    float fv;
    long nxf = (long)ny * (nx + 1) * ngf;
    long f;
    int i, j;
    char head[96];
    memcpy(&fv, &vb, sizeof(fv));
    if (key < nxf) {
        f = key / ngf;
        i = (int)(f % (nx + 1));
        j = (int)(f / (nx + 1));
    } else {
        f = (key - nxf) / ngf;
        i = (int)(f % nx);
        j = (int)(f / nx);
    }
    *x = x_min + (double)(i)*h;
    *y = y_min + (double)(j)*h;
    if (kind < 1 || kind > 7) kind = 0;
    snprintf(head, sizeof(head), what[kind], (double)fv);
    snprintf(msg, cap, "%s\nAt x = %f, y = %f", head, *x, *y);
}

/* tile layouts (bit 0: one wave per tile, bit 1: split tiles) a problem may
 * use: those the system's kernel is built for, narrowed by the layout bits
 * of its options */
int hype_layouts(const hype_problem *p, int built)
{
    if (p->options & HYPE_OPT_ONE_WAVE_TILES) return built & 1;
    if (p->options & HYPE_OPT_SPLIT_TILES) return built & 2;
    return built;
}

void hype_fill_inst(hype_inst *I, const hype_problem *p, int nsy)
{
    int s;
    memset(I, 0, sizeof(*I));
    I->nx = p->nx;
    I->ny = p->ny;
    for (s = 0; s < 4; ++s) I->bc[s] = p->bc[s];
    I->nsx = hype_strips(p->nx);
    I->nsy = nsy;
    I->h = (p->x_max - p->x_min) / (double)(p->nx); /* hype.c:57 */
    I->r1_h = 1. / (I->h);                          /* rhs_function_prim.c:26 */
    I->cfl = p->cfl;
    I->x_min = p->x_min;
    I->y_min = p->y_min;
    I->g1 = p->gamma1;
    I->g2 = p->gamma2;
    I->p1 = p->pi1;
    I->p2 = p->pi2;
    I->quirks = p->quirks;
    I->keybits = hype_key_bits(p->nx, p->ny);
}

/* Cell-centre coordinates.  HYPE_COORD_EXACT: x_min + h/2 + i h.
 * HYPE_COORD_DMDA: the coordinates DMDASetUniformCoordinates(x_min + h/2,
 * x_max + h/2, ...) gives the reference (hype.c:132-135): spacing h on a
 * periodic axis, span/(n - 1) otherwise. */
double hype_cell_coord(const hype_problem *p, int axis, int i, int coord_mode)
{
    const double h = (p->x_max - p->x_min) / (double)p->nx; /* hype.c:57 */
    double d = h;
    if (axis == 0) {
        if (coord_mode == HYPE_COORD_DMDA && p->bc[0] != HYPE_BC_PERIODIC && p->bc[1] != HYPE_BC_PERIODIC)
            d = (p->x_max - p->x_min) / (double)(p->nx - 1);
        return p->x_min + 0.5 * h + (double)i * d;
    }
    if (coord_mode == HYPE_COORD_DMDA && p->bc[2] != HYPE_BC_PERIODIC && p->bc[3] != HYPE_BC_PERIODIC)
        d = ((p->y_max + 0.5 * h) - (p->y_min + 0.5 * h)) / (double)(p->ny - 1);
    return p->y_min + 0.5 * h + (double)i * d;
}

/* XCD-aware order of a row-major tile list.  Workgroups of a launch are
 * dispatched round-robin over the device's XCDs (workgroup i runs on XCD
 * i mod nxcd), and every XCD has its own L2.  Giving each XCD a contiguous
 * band of the row-major list (list[i] = band[i mod nxcd][i / nxcd]) keeps
 * vertically adjacent tiles, which read each other's three halo rows, on
 * the same L2. */
void hype_xcd_order_into(const hype_tile *src, hype_tile *dst, int n, int nxcd)
{
    int x, i, start = 0;
    if (nxcd < 2 || n <= nxcd) {
        memcpy(dst, src, (size_t)n * sizeof(hype_tile));
        return;
    }
    for (x = 0; x < nxcd; ++x) {
        const int cnt = n / nxcd + (x < n % nxcd ? 1 : 0); /* slots i = x, x + nxcd, ... below n */
        for (i = 0; i < cnt; ++i) dst[x + i * nxcd] = src[start + i];
        start += cnt;
    }
}

int hype_xcd_order(hype_tile *tiles, int n, int nxcd)
{
    hype_tile *tmp;
    if (nxcd < 2 || n <= nxcd) return HYPE_OK;
    tmp = (hype_tile *)malloc((size_t)n * sizeof(hype_tile));
    if (!tmp) return HYPE_E_NOMEM;
    memcpy(tmp, tiles, (size_t)n * sizeof(hype_tile));
    hype_xcd_order_into(tmp, tiles, n, nxcd);
    free(tmp);
    return HYPE_OK;
}

int hype_pick_layout(const hype_problem *p, int *pair, int *nsy, int *slots)
{
    const int modes = hype_layouts(p, hyped_pair_modes(p->system));
    int ncu = hyped_num_cus(), pr, found = 0;
    double best = 1e300;
    if (ncu < 1) ncu = 256;
    for (pr = 0; pr <= 1; ++pr) {
        hype_variant v;
        int sl, occ, ns;
        double cost;
        if (!(modes & (1 << pr))) continue;
        v.system = p->system;
        v.prim = p->recons_prim;
        v.order = p->order;
        v.fma = (p->options & HYPE_OPT_FMA) != 0;
        v.pair = pr;
        v.inlet = p->bc[0] == HYPE_BC_INLET;
        sl = ncu * hyped_blocks_per_cu(&v) * hyped_block_waves(&v); /* resident waves */
        if (sl < 1) sl = 4 * ncu;
        occ = (sl + 4 * ncu - 1) / (4 * ncu);
        ns = hype_pick_segments(p->nx, p->ny, 4 * ncu, occ, 0, pr);
        cost = hype_tile_cost((long)hype_strips(p->nx) * ns * hyped_block_waves(&v),
                              hype_tile_rows((p->ny + ns - 1) / ns, pr), 4 * ncu, occ);
        if (cost < best - 1e-9) {
            best = cost;
            *pair = pr;
            *nsy = ns;
            if (slots) *slots = sl;
            found = 1;
        }
    }
    return found ? HYPE_OK : HYPE_E_UNSUPPORTED;
}
