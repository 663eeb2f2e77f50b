/*
 * orc_rhs.c -- the RHS callback restated (oracle; test infrastructure only).
 *
 * Follows the pass structure of RHSFunctionPrim (rhs_function_prim.c:12-626)
 * and RHSFunction (rhs_function.c:12-625) on one rank:
 *   1 primitive pass (prim path only)         rhs_function_prim.c:40
 *   2 ghosted local copy, periodic wrap       rhs_function_prim.c:44-45 (DMGlobalToLocal)
 *   3 boundary fill, x-pass then y-pass       rhs_function_prim.c:75-315
 *   4 reconstruction of the ring-extended block into face scratch u_bnd,
 *     admissibility check + fallback, volume NCP   :321-481
 *   5 x-face Riemann pass                     :485-533
 *   6 y-face Riemann pass                     :537-585
 *   7 assembly                                :589-602
 *   8 dt = CFL*h/(2*s_max)                    :620-623
 * OpenMP pragmas parallelise the cell and face loops of the timed build; the
 * parity build is compiled without OpenMP.  s_max is a max and the recorded
 * error is the smallest face key, so results do not depend on threading.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

void orc_basis_grad(double x, double y, int n, double *gx, double *gy);

typedef struct {
    orc_problem P;
    int nv, nDOF, ngf, ngv, nnode;
    double xGP_Face[2], wGP_Face[2];
    double xGP_Vol[4], yGP_Vol[4], wGP_Vol[4];
    double x_node[8], y_node[8];
    double phiFace[4][2][ORC_MAXDOF];
    double phiNode[8][ORC_MAXDOF];
    double phiVol[4][ORC_MAXDOF], gphiVx[4][ORC_MAXDOF], gphiVy[4][ORC_MAXDOF];
    double *W;     /* nx*ny*nv primitive pass */
    double *loc;   /* (ny+6)*(nx+6)*nv ghosted local array */
    double *ubnd;  /* (ny+2)*(nx+2)*nv*4*ngf */
    double *F, *D; /* ny*(nx+1)*nv */
    double *G, *E; /* (ny+1)*nx*nv */
} orc_ctx;

/* velocity/momentum components negated by a reflective wall
 * (rhs_function_prim.c:101-129, 169-197, 230-258, 283-311) */
static int refl_x(int system, int k)
{
    if (system == ORC_EULER) return k == 0 ? 1 : -1;
    if (system == ORC_KAPILA) return k == 0 ? 2 : -1;
    return k == 0 ? 1 : (k == 1 ? 5 : -1);
}
static int refl_y(int system, int k)
{
    if (system == ORC_EULER) return k == 0 ? 2 : -1;
    if (system == ORC_KAPILA) return k == 0 ? 3 : -1;
    return k == 0 ? 2 : (k == 1 ? 6 : -1);
}

static void ctx_tables(orc_ctx *C)
{
    const orc_problem *P = &C->P;
    int q, k;
    C->nv = orc_nvar(P->system);
    C->nDOF = (P->N + 1) * (P->N + 2) / 2;                 /* hype.c:56 */
    if (P->N < 2) { C->ngf = 1; C->ngv = 1; }              /* hype.c:59-66 */
    else { C->ngf = 2; C->ngv = 4; }
    C->nnode = 4 * C->ngf;                                 /* hype.c:68 */
    if (P->N < 2) {                                        /* hype.c:203-211 */
        C->xGP_Face[0] = 0.0; C->wGP_Face[0] = 1.0;
        C->xGP_Vol[0] = 0.0; C->yGP_Vol[0] = 0.0; C->wGP_Vol[0] = 1.0;
        C->x_node[0] = 0.5;  C->y_node[0] = 0.0;
        C->x_node[1] = -0.5; C->y_node[1] = 0.0;
        C->x_node[2] = 0.0;  C->y_node[2] = 0.5;
        C->x_node[3] = 0.0;  C->y_node[3] = -0.5;
    } else {                                               /* hype.c:213-230 */
        const double g = 0.28867513459481287;
        C->xGP_Face[0] = -g; C->wGP_Face[0] = 0.5;
        C->xGP_Face[1] = g;  C->wGP_Face[1] = 0.5;
        C->xGP_Vol[0] = -g; C->yGP_Vol[0] = -g; C->wGP_Vol[0] = 0.25;
        C->xGP_Vol[1] = -g; C->yGP_Vol[1] = g;  C->wGP_Vol[1] = 0.25;
        C->xGP_Vol[2] = g;  C->yGP_Vol[2] = -g; C->wGP_Vol[2] = 0.25;
        C->xGP_Vol[3] = g;  C->yGP_Vol[3] = g;  C->wGP_Vol[3] = 0.25;
        C->x_node[0] = 0.5;  C->y_node[0] = -g;
        C->x_node[1] = 0.5;  C->y_node[1] = g;
        C->x_node[2] = -0.5; C->y_node[2] = -g;
        C->x_node[3] = -0.5; C->y_node[3] = g;
        C->x_node[4] = -g;   C->y_node[4] = 0.5;
        C->x_node[5] = g;    C->y_node[5] = 0.5;
        C->x_node[6] = -g;   C->y_node[6] = -0.5;
        C->x_node[7] = g;    C->y_node[7] = -0.5;
    }
    for (q = 0; q < C->ngf; ++q)                            /* hype.c:255-286 */
        for (k = 0; k < C->nDOF; ++k) {
            C->phiFace[0][q][k] = orc_basis(-0.5, C->xGP_Face[q], k);
            C->phiFace[1][q][k] = orc_basis(0.5, C->xGP_Face[q], k);
            C->phiFace[2][q][k] = orc_basis(C->xGP_Face[q], -0.5, k);
            C->phiFace[3][q][k] = orc_basis(C->xGP_Face[q], 0.5, k);
        }
    for (q = 0; q < C->nnode; ++q)                          /* hype.c:290-294 */
        for (k = 0; k < C->nDOF; ++k)
            C->phiNode[q][k] = orc_basis(C->x_node[q], C->y_node[q], k);
    for (q = 0; q < C->ngv; ++q)                            /* hype.c:298-305 */
        for (k = 0; k < C->nDOF; ++k) {
            C->phiVol[q][k] = orc_basis(C->xGP_Vol[q], C->yGP_Vol[q], k);
            orc_basis_grad(C->xGP_Vol[q], C->yGP_Vol[q], k, &C->gphiVx[q][k],
                           &C->gphiVy[q][k]);
        }
}

static int check_problem(const orc_problem *P)
{
    int s;
    if (P->system < 0 || P->system > 2) return ORC_E_INVALID;
    if (P->N < 0 || P->N > 3) return ORC_E_INVALID;
    if (P->nx < 3 || P->ny < 3) return ORC_E_INVALID;
    if (!(P->h > 0.0)) return ORC_E_INVALID;
    for (s = 0; s < 4; ++s) {
        if (P->bc[s] == ORC_BC_WALL)
            return ORC_E_UNSUPPORTED; /* no implementation (hype.h:76) */
        /* the inlet exists on the left boundary of the Kapila system only
         * (rhs_function_prim.c:135-144; InletBC in kapila_5eqn.c alone) */
        if (P->bc[s] == ORC_BC_INLET && (s != 0 || P->system != ORC_KAPILA)) return ORC_E_UNSUPPORTED;
        if (P->bc[s] < 0 || P->bc[s] > 4) return ORC_E_INVALID;
    }
    return ORC_OK;
}

void *orc_ctx_create(const orc_problem *P)
{
    if (check_problem(P) != ORC_OK) return NULL;
    orc_ctx *C = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!C) return NULL;
    C->P = *P;
    ctx_tables(C);
    size_t nx = (size_t)P->nx, ny = (size_t)P->ny, nv = (size_t)C->nv;
    C->W = (double *)malloc(nx * ny * nv * sizeof(double));
    C->loc = (double *)calloc((nx + 6) * (ny + 6) * nv, sizeof(double));
    C->ubnd = (double *)malloc((nx + 2) * (ny + 2) * nv * 4 * C->ngf * sizeof(double));
    C->F = (double *)malloc(ny * (nx + 1) * nv * sizeof(double));
    C->D = (double *)malloc(ny * (nx + 1) * nv * sizeof(double));
    C->G = (double *)malloc((ny + 1) * nx * nv * sizeof(double));
    C->E = (double *)malloc((ny + 1) * nx * nv * sizeof(double));
    if (!C->W || !C->loc || !C->ubnd || !C->F || !C->D || !C->G || !C->E) {
        free(C->W); free(C->loc); free(C->ubnd); free(C->F); free(C->D);
        free(C->G); free(C->E); free(C);
        return NULL;
    }
    return C;
}

void orc_ctx_destroy(void *vc)
{
    orc_ctx *C = (orc_ctx *)vc;
    if (!C) return;
    free(C->W); free(C->loc); free(C->ubnd); free(C->F); free(C->D);
    free(C->G); free(C->E); free(C);
}

/* ghosted local array, indices j,i in [-3, n+2] */
#define LOC(C, j, i) ((C)->loc + ((size_t)((j) + 3) * ((C)->P.nx + 6) + (size_t)((i) + 3)) * (C)->nv)
#define UBND(C, lj, li, c, f, q) \
    ((C)->ubnd[((((size_t)(lj) * ((C)->P.nx + 2) + (size_t)(li)) * (C)->nv + (c)) * 4 + (f)) * (C)->ngf + (q)])

static void fill_local(orc_ctx *C, const double *src)
{
    const orc_problem *P = &C->P;
    const int nx = P->nx, ny = P->ny, nv = C->nv;
    int i, j, c, k;
    /* DMGlobalToLocal: interior copy; periodic axes wrap (hype.c:98-106) */
    int xper = (P->bc[0] == ORC_BC_PERIODIC || P->bc[1] == ORC_BC_PERIODIC);
    int yper = (P->bc[2] == ORC_BC_PERIODIC || P->bc[3] == ORC_BC_PERIODIC);
    for (j = -3; j < ny + 3; ++j)
        for (i = -3; i < nx + 3; ++i) {
            int si = i, sj = j, ok = 1;
            if (i < 0 || i >= nx) { if (xper) si = (i + nx) % nx; else ok = 0; }
            if (j < 0 || j >= ny) { if (yper) sj = (j + ny) % ny; else ok = 0; }
            double *d = LOC(C, j, i);
            if (ok) memcpy(d, src + ((size_t)sj * nx + si) * nv, nv * sizeof(double));
            else memset(d, 0, nv * sizeof(double)); /* DMCreateLocalVector zeroes */
        }
    /* x-pass over the whole ghosted range (rhs_function_prim.c:75-202) */
    for (j = -3; j < ny + 3; ++j)
        for (i = -3; i < nx + 3; ++i) {
            int side = -1, irhs = 0;
            if (i < 0) { side = 0; }
            if (i >= nx) { side = 1; }
            if (side < 0) continue;
            int bc = P->bc[side];
            if (bc == ORC_BC_INLET) { /* rhs_function_prim.c:135-144 / rhs_function.c:135-144 */
                double V[ORC_MAXVAR];
                const double y_loc = P->y_min + ((double)(j) + 0.5) * P->h;
                double *d = LOC(C, j, i);
                orc_inlet_prim(P, y_loc, V);
                if (P->prim) for (c = 0; c < nv; ++c) d[c] = V[c];
                else orc_prim2cons(P, V, d);
                continue;
            }
            if (bc == ORC_BC_TRANSMISSIVE) irhs = side == 0 ? 0 : nx - 1;
            else if (bc == ORC_BC_REFLECTIVE) irhs = side == 0 ? -1 - i : 2 * (nx - 1) - i + 1;
            else continue;
            double *d = LOC(C, j, i);
            const double *s = LOC(C, j, irhs);
            for (c = 0; c < nv; ++c) d[c] = s[c];
            if (bc == ORC_BC_REFLECTIVE)
                for (k = 0; k < 2; ++k) {
                    int cc = refl_x(P->system, k);
                    if (cc >= 0) d[cc] = -d[cc];
                }
        }
    /* y-pass over the whole ghosted range (rhs_function_prim.c:204-315) */
    for (j = -3; j < ny + 3; ++j)
        for (i = -3; i < nx + 3; ++i) {
            int side = -1, jrhs = 0;
            if (j < 0) side = 2;
            if (j >= ny) side = 3;
            if (side < 0) continue;
            int bc = P->bc[side];
            if (bc == ORC_BC_TRANSMISSIVE) jrhs = side == 2 ? 0 : ny - 1;
            else if (bc == ORC_BC_REFLECTIVE) jrhs = side == 2 ? -1 - j : 2 * (ny - 1) - j + 1;
            else continue;
            double *d = LOC(C, j, i);
            const double *s = LOC(C, jrhs, i);
            for (c = 0; c < nv; ++c) d[c] = s[c];
            if (bc == ORC_BC_REFLECTIVE)
                for (k = 0; k < 2; ++k) {
                    int cc = refl_y(P->system, k);
                    if (cc >= 0) d[cc] = -d[cc];
                }
        }
}

int orc_rhs_ctx(void *vc, double t, const double *U, double *RHS,
                double *dt_out, double *W, unsigned char *fallback,
                orc_error *err)
{
    orc_ctx *C = (orc_ctx *)vc;
    const orc_problem *P = &C->P;
    const int nx = P->nx, ny = P->ny, nv = C->nv, nDOF = C->nDOF;
    const int ngf = C->ngf, ngv = C->ngv, nnode = C->nnode;
    const double r1_h = 1. / (P->h);
    double s_max_c = 0.0;
    long best_key = LONG_MAX;
    int best_kind = 0;
    double best_x = 0, best_y = 0, best_val = 0;
    int i, j;
    (void)t;

    if (err) memset(err, 0, sizeof(*err));

    /* 1-3: primitive pass, ghosted copy, boundary fill */
    if (P->prim) {
        long n = (long)nx * ny, m;
#pragma omp parallel for schedule(static)
        for (m = 0; m < n; ++m) orc_cons2prim(P, U + m * nv, C->W + m * nv);
        fill_local(C, C->W);
        if (W) memcpy(W, C->W, (size_t)n * nv * sizeof(double));
    } else {
        fill_local(C, U);
    }

    /* 4: reconstruction over the ring-extended block */
#pragma omp parallel for schedule(static) private(i)
    for (j = -1; j < ny + 1; ++j) {
        for (i = -1; i < nx + 1; ++i) {
            double wx[5], wy[5], wxy[5], coeffs[ORC_MAXDOF];
            double sol[ORC_MAXDOF][ORC_MAXVAR], VNode[8][ORC_MAXVAR];
            double Vgp[4][ORC_MAXVAR], gVx[4][ORC_MAXVAR], gVy[4][ORC_MAXVAR];
            int lj = j + 1, li = i + 1, c, k, q, f, PAD = 1;
            for (c = 0; c < nv; ++c) {
                wx[0] = LOC(C, j, i - 2)[c];
                wx[1] = LOC(C, j, i - 1)[c];
                wx[2] = LOC(C, j, i)[c];
                wx[3] = LOC(C, j, i + 1)[c];
                wx[4] = LOC(C, j, i + 2)[c];
                wy[0] = LOC(C, j - 2, i)[c];
                wy[1] = LOC(C, j - 1, i)[c];
                wy[2] = LOC(C, j, i)[c];
                wy[3] = LOC(C, j + 1, i)[c];
                wy[4] = LOC(C, j + 2, i)[c];
                wxy[0] = LOC(C, j, i)[c];
                wxy[1] = LOC(C, j + 1, i + 1)[c];
                wxy[2] = LOC(C, j - 1, i + 1)[c];
                wxy[3] = LOC(C, j + 1, i - 1)[c];
                wxy[4] = LOC(C, j - 1, i - 1)[c];
                orc_reconstruct(wx, wy, wxy, P->N, coeffs);
                for (k = 0; k < nDOF; ++k) sol[k][c] = coeffs[k];
                for (q = 0; q < nnode; ++q) {
                    VNode[q][c] = 0.0;
                    for (k = 0; k < nDOF; ++k) VNode[q][c] += coeffs[k] * C->phiNode[q][k];
                }
            }
            for (q = 0; q < nnode; ++q) {                   /* :393-404 */
                PAD = P->prim ? orc_pad_prim(P, VNode[q]) : orc_pad(P, VNode[q]);
                if (!PAD) break;
            }
            if (!PAD) {                                     /* :407-421 / rhs_function.c:409-421 */
                for (c = 0; c < nv; ++c) {
                    const double *w0 = LOC(C, j, i);
                    sol[0][c] = w0[c];
                    if (nDOF > 1) {
                        sol[1][c] = orc_minmod(LOC(C, j, i + 1)[c] - w0[c], w0[c] - LOC(C, j, i - 1)[c]);
                        sol[2][c] = orc_minmod(LOC(C, j + 1, i)[c] - w0[c], w0[c] - LOC(C, j - 1, i)[c]);
                    }
                    for (k = P->prim ? 3 : 1; k < nDOF; ++k) sol[k][c] = 0.0;
                }
            }
            if (fallback && i >= 0 && i < nx && j >= 0 && j < ny)
                fallback[(size_t)j * nx + i] = (unsigned char)(!PAD);
            for (c = 0; c < nv; ++c) {                      /* :425-456 */
                for (f = 0; f < 4; ++f)
                    for (q = 0; q < ngf; ++q) {
                        double value = 0.0;
                        for (k = 0; k < nDOF; ++k) value += sol[k][c] * C->phiFace[f][q][k];
                        UBND(C, lj, li, c, f, q) = value;
                    }
                for (q = 0; q < ngv; ++q) {
                    double value = 0.0, gx = 0.0, gy = 0.0;
                    for (k = 0; k < nDOF; ++k) {
                        value += sol[k][c] * C->phiVol[q][k];
                        gx += sol[k][c] * C->gphiVx[q][k];
                        gy += sol[k][c] * C->gphiVy[q][k];
                    }
                    Vgp[q][c] = value;
                    gVx[q][c] = r1_h * gx;
                    gVy[q][c] = r1_h * gy;
                }
            }
            if (i >= 0 && i < nx && j >= 0 && j < ny) {     /* :461-479 */
                double *r = RHS + ((size_t)j * nx + i) * nv;
                double B[ORC_MAXVAR];
                for (c = 0; c < nv; ++c) r[c] = 0.0;
                for (q = 0; q < ngv; ++q) {
                    if (P->prim) orc_ncp_prim(P, Vgp[q], gVx[q], gVy[q], B);
                    else orc_ncp(P, Vgp[q], gVx[q], gVy[q], B);
                    for (c = 0; c < nv; ++c) r[c] += -C->wGP_Vol[q] * B[c];
                }
            }
        }
    }

    /* 5: x-faces (rhs_function_prim.c:485-533) */
#pragma omp parallel private(i)
    {
        double smax_t = 0.0, bx = 0, by = 0, bv = 0;
        long bkey = LONG_MAX;
        int bkind = 0;
#pragma omp for schedule(static)
        for (j = 0; j < ny; ++j) {
            for (i = 0; i < nx + 1; ++i) {
                int lj = j + 1, li = i + 1, c, q;
                double x_loc = P->x_min + (double)(i) * P->h;
                double y_loc = P->y_min + (double)(j) * P->h;
                double Flux[ORC_MAXVAR], NCP[ORC_MAXVAR], VL[ORC_MAXVAR], VR[ORC_MAXVAR];
                double Fc[ORC_MAXVAR], Fl[ORC_MAXVAR];
                for (c = 0; c < nv; ++c) { Flux[c] = 0.0; NCP[c] = 0.0; }
                for (q = 0; q < ngf; ++q) {
                    int bad = 0;
                    double bval = 0.0;
                    for (c = 0; c < nv; ++c) {
                        VL[c] = UBND(C, lj, li - 1, c, 1, q);
                        VR[c] = UBND(C, lj, li, c, 0, q);
                    }
                    double s_c = orc_riemann_x(P, VL, VR, 1.0, 0.0, Fc, Fl, &bad, &bval);
                    if (s_c > smax_t) smax_t = s_c;
                    if (bad) {
                        long key = ((long)j * (nx + 1) + i) * ngf + q;
                        if (key < bkey) { bkey = key; bkind = bad; bx = x_loc; by = y_loc; bv = bval; }
                    }
                    for (c = 0; c < nv; ++c) {
                        Flux[c] += C->wGP_Face[q] * Fc[c];
                        NCP[c] += C->wGP_Face[q] * Fl[c];
                    }
                }
                for (c = 0; c < nv; ++c) {
                    C->F[((size_t)j * (nx + 1) + i) * nv + c] = Flux[c];
                    C->D[((size_t)j * (nx + 1) + i) * nv + c] = NCP[c];
                }
            }
        }
        /* 6: y-faces (rhs_function_prim.c:537-585) */
#pragma omp for schedule(static)
        for (j = 0; j < ny + 1; ++j) {
            for (i = 0; i < nx; ++i) {
                int lj = j + 1, li = i + 1, c, q;
                double x_loc = P->x_min + (double)(i) * P->h;
                double y_loc = P->y_min + (double)(j) * P->h;
                double Flux[ORC_MAXVAR], NCP[ORC_MAXVAR], VL[ORC_MAXVAR], VR[ORC_MAXVAR];
                double Fc[ORC_MAXVAR], Fl[ORC_MAXVAR];
                for (c = 0; c < nv; ++c) { Flux[c] = 0.0; NCP[c] = 0.0; }
                for (q = 0; q < ngf; ++q) {
                    int bad = 0;
                    double bval = 0.0;
                    for (c = 0; c < nv; ++c) {
                        VL[c] = UBND(C, lj - 1, li, c, 3, q);
                        VR[c] = UBND(C, lj, li, c, 2, q);
                    }
                    double s_c = orc_riemann_x(P, VL, VR, 0.0, 1.0, Fc, Fl, &bad, &bval);
                    if (s_c > smax_t) smax_t = s_c;
                    if (bad) {
                        long key = (long)ny * (nx + 1) * ngf + ((long)j * nx + i) * ngf + q;
                        if (key < bkey) { bkey = key; bkind = bad; bx = x_loc; by = y_loc; bv = bval; }
                    }
                    for (c = 0; c < nv; ++c) {
                        Flux[c] += C->wGP_Face[q] * Fc[c];
                        NCP[c] += C->wGP_Face[q] * Fl[c];
                    }
                }
                for (c = 0; c < nv; ++c) {
                    C->G[((size_t)j * nx + i) * nv + c] = Flux[c];
                    C->E[((size_t)j * nx + i) * nv + c] = NCP[c];
                }
            }
        }
#pragma omp critical
        {
            if (smax_t > s_max_c) s_max_c = smax_t;
            if (bkey < best_key) {
                best_key = bkey; best_kind = bkind; best_x = bx; best_y = by; best_val = bv;
            }
        }
    }

    /* 7: assembly (rhs_function_prim.c:589-602) */
#pragma omp parallel for schedule(static) private(i)
    for (j = 0; j < ny; ++j)
        for (i = 0; i < nx; ++i) {
            double *r = RHS + ((size_t)j * nx + i) * nv;
            const double *F0 = C->F + ((size_t)j * (nx + 1) + i) * nv;
            const double *F1 = F0 + nv;
            const double *D0 = C->D + ((size_t)j * (nx + 1) + i) * nv;
            const double *D1 = D0 + nv;
            const double *G0 = C->G + ((size_t)j * nx + i) * nv;
            const double *G1 = C->G + ((size_t)(j + 1) * nx + i) * nv;
            const double *E0 = C->E + ((size_t)j * nx + i) * nv;
            const double *E1 = C->E + ((size_t)(j + 1) * nx + i) * nv;
            int c;
            for (c = 0; c < nv; ++c)
                r[c] += -r1_h * (F1[c] - F0[c] + 0.5 * (D1[c] + D0[c])) -
                        r1_h * (G1[c] - G0[c] + 0.5 * (E1[c] + E0[c]));
        }

    /* 8: dt (rhs_function_prim.c:620-623) */
    if (dt_out) *dt_out = (P->cfl * P->h) / (2.0 * s_max_c);
    if (best_kind) {
        if (err) {
            err->code = ORC_E_INADMISSIBLE;
            err->kind = best_kind;
            err->x = best_x;
            err->y = best_y;
            err->value = best_val;
            err->index = best_key;
        }
        return ORC_E_INADMISSIBLE;
    }
    return ORC_OK;
}

int orc_rhs(const orc_problem *P, double t, const double *U, double *RHS,
            double *dt_out, double *W, unsigned char *fallback,
            orc_error *err)
{
    int rc = check_problem(P);
    if (rc != ORC_OK) return rc;
    void *C = orc_ctx_create(P);
    if (!C) return ORC_E_INVALID;
    rc = orc_rhs_ctx(C, t, U, RHS, dt_out, W, fallback, err);
    orc_ctx_destroy(C);
    return rc;
}
