/* common.c -- problem validation, tiling and error decoding (host side). */
#include "common.h"

#include <stdio.h>
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
    if (p->order < 0 || p->order > 2) return HYPE_E_UNSUPPORTED;
    if (p->nx < 3 || p->ny < 3) return HYPE_E_INVALID;
    /* face keys of the error word hold 28 bits */
    if ((double)(p->nx + 1) * p->ny * 2.0 * 2.0 >= (double)(1 << 28)) return HYPE_E_UNSUPPORTED;
    if (!(p->x_max > p->x_min) || !(p->y_max > p->y_min)) return HYPE_E_INVALID;
    if (!(p->cfl > 0.0)) return HYPE_E_INVALID;
    for (s = 0; s < 4; ++s) {
        if (p->bc[s] == HYPE_BC_WALL || p->bc[s] == HYPE_BC_INLET) return HYPE_E_UNSUPPORTED;
        if (p->bc[s] != HYPE_BC_PERIODIC && p->bc[s] != HYPE_BC_REFLECTIVE &&
            p->bc[s] != HYPE_BC_TRANSMISSIVE)
            return HYPE_E_INVALID;
    }
    if (nvar) *nvar = nv;
    return HYPE_OK;
}

int hype_count_tiles(int nx, int ny, int hrows)
{
    int tx = (nx + HYPE_TILE_W - 1) / HYPE_TILE_W;
    int ty = (ny + hrows - 1) / hrows;
    return tx * ty;
}

int hype_make_tiles(int inst, int nx, int ny, int hrows, hype_tile *tiles, int cap)
{
    int n = 0, x0, y0;
    for (y0 = 0; y0 < ny; y0 += hrows)
        for (x0 = 0; x0 < nx; x0 += HYPE_TILE_W) {
            if (n >= cap) return -1;
            tiles[n].inst = inst;
            tiles[n].x0 = x0;
            tiles[n].y0 = y0;
            tiles[n].flags = (n == 0) ? 1 : 0;
            ++n;
        }
    return n;
}

/* Rows per tile.  A tile costs (hrows+2) reconstructed rows for hrows owned
 * ones, so tall tiles waste less; but the grid must offer enough tiles to
 * keep every SIMD of the device busy (at least ~4 waves per CU). */
int hype_default_rows(int nx, int ny, int system, int ncu)
{
    int tx = (nx + HYPE_TILE_W - 1) / HYPE_TILE_W;
    int want = (ncu > 0 ? ncu : 256) * 4;
    int h = 64;
    (void)system;
    while (h > 8 && (long)tx * ((ny + h - 1) / h) < want) h /= 2;
    if (h > ny) h = ny;
    return h;
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
    long key = (long)((word >> HYPE_ERR_FACE_SHIFT) & 0xFFFFFFFull);
    int kind = (int)((word >> HYPE_ERR_KIND_SHIFT) & 0xF);
    unsigned vb = (unsigned)(word & 0x1FFFFFFFull) << 3;
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

void hype_fill_inst(hype_inst *I, const hype_problem *p, int hrows)
{
    int s;
    memset(I, 0, sizeof(*I));
    I->nx = p->nx;
    I->ny = p->ny;
    for (s = 0; s < 4; ++s) I->bc[s] = p->bc[s];
    I->hrows = hrows;
    I->ntiles_x = (p->nx + HYPE_TILE_W - 1) / HYPE_TILE_W;
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
}
